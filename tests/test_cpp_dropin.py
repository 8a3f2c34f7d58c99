"""The C++ drop-in header (include/collision_detection.h) used the way the
reference's own test mains use the reference classes.

The mains compile on the CPU (non-GPU suite); running them needs the GPU.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
PKG = os.path.join(ROOT, "collision-detection_amd")
MAINS = ["test1", "binTest", "nSquared", "pipeline"]


def _build(name, tmp, defines=()):
    exe = os.path.join(tmp, name)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT}/include", *defines, os.path.join(CPP, name + ".cpp"),
           f"-L{PKG}", "-lcdx", f"-Wl,-rpath,{PKG}", "-o", exe]
    subprocess.check_call(cmd)
    return exe


@pytest.fixture(scope="module")
def exes(tmp_path_factory, cdx):
    assert os.path.exists(cdx.LIB_PATH), "libcdx.so not built"
    tmp = str(tmp_path_factory.mktemp("cpp"))
    return {m: _build(m, tmp) for m in MAINS}


def test_mains_compile_against_header(exes):
    for m in MAINS:
        assert os.path.exists(exes[m])


def test_header_compiles_in_float_mode(tmp_path, cdx):
    _build("test1", str(tmp_path), defines=("-DCDX_REAL_FLOAT",))


def _run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    return r


@pytest.mark.gpu
def test_test1_main(exes):
    for n in (10, 1000):
        r = _run(exes["test1"], n)
        assert r.returncode == 0 and "Test Passed!" in r.stdout


@pytest.mark.gpu
def test_bintest_main(exes):
    for args in ((10, 10), (20, 7)):
        r = _run(exes["binTest"], *args)
        assert r.returncode == 0 and "TEST PASSED" in r.stdout


@pytest.mark.gpu
def test_nsquared_main(exes):
    r = _run(exes["nSquared"])
    assert r.returncode == 0 and "TEST PASSED" in r.stdout


def _read_dump(path):
    out = []
    with open(path, "rb") as f:
        spec = [("T", np.int32, 1), ("A", np.float64, 3), ("B", np.float64, 3), ("R", np.float64, 4),
                ("ID", np.uint32, 1), ("pos", np.float64, 3), ("rot", np.float64, 4), ("active", np.uint8, 1),
                ("aabb", np.float64, 3), ("aabb_shifted", np.float64, 3), ("pairs", np.int64, 1),
                ("pairs_permuted", np.int64, 1), ("norm", np.float64, 3), ("cpta", np.float64, 3),
                ("cptb", np.float64, 3), ("dpth", np.float64, 1), ("bids", np.int32, 2),
                ("st_active", np.uint32, 1), ("st_norm", np.float64, 3), ("st_cpta", np.float64, 3),
                ("st_cptb", np.float64, 3), ("st_dpth", np.float64, 1), ("st_bids", np.int32, 2)]
        d = {}
        for name, dt, k in spec:
            n = int(np.frombuffer(f.read(8), np.uint64)[0])
            a = np.frombuffer(f.read(n * np.dtype(dt).itemsize * k), dt).copy()
            d[name] = a.reshape(n, k) if k > 1 else a
    return d


@pytest.mark.gpu
def test_pipeline_main_vs_oracle(exes, tmp_path, oracle, scenes):
    dump = str(tmp_path / "pipeline.bin")
    r = _run(exes["pipeline"], dump, 6000)
    assert r.returncode == 0
    d = _read_dump(dump)
    n = len(d["T"])
    s = scenes.Scene(d["T"], d["A"], d["B"], np.zeros((n, 3)), d["R"], d["ID"], d["pos"], d["rot"], d["active"],
                     (12, 12, 12))
    ab = oracle.aabb(s)
    np.testing.assert_array_equal(d["aabb"], ab)
    bp = oracle.broadphase(ab, (12, 12, 12))
    np.testing.assert_array_equal(d["aabb_shifted"], bp.aabb_shifted)
    np.testing.assert_array_equal(d["pairs"], bp.pairs)
    npr = oracle.narrowphase(s, bp.pairs, envelope=0.0005)
    np.testing.assert_array_equal(d["pairs_permuted"], npr.pairs_permuted)
    np.testing.assert_array_equal(d["bids"], npr.bids)
    np.testing.assert_allclose(d["norm"], npr.norm, atol=1e-12, rtol=0)
    np.testing.assert_allclose(d["cpta"], npr.cpta, atol=1e-12, rtol=0)
    np.testing.assert_allclose(d["cptb"], npr.cptb, atol=1e-12, rtol=0)
    np.testing.assert_allclose(d["dpth"], npr.dpth, atol=1e-12, rtol=0)


@pytest.mark.gpu
def test_host_mpr_store_main_vs_oracle(exes, tmp_path, oracle, scenes):
    """Narrowphase::host_MPR_Store (Narrowphase.cu:464-504): per-pair arrays
    at the pair index; contacts clear contact_active, every other entry
    keeps the caller's value."""
    dump = str(tmp_path / "pipeline.bin")
    r = _run(exes["pipeline"], dump, 6000)
    assert r.returncode == 0
    d = _read_dump(dump)
    n = len(d["T"])
    s = scenes.Scene(d["T"], d["A"], d["B"], np.zeros((n, 3)), d["R"], d["ID"], d["pos"], d["rot"], d["active"],
                     (12, 12, 12))
    pairs = d["pairs"]
    npr = oracle.narrowphase(s, pairs, envelope=0.0005)
    hit = npr.hit.astype(bool)
    assert hit.sum() == npr.count > 100 and (~hit).sum() > 100
    np.testing.assert_array_equal(d["st_active"], np.where(hit, 0, 1).astype(np.uint32))
    for k, ref in (("st_norm", npr.norm), ("st_cpta", npr.cpta), ("st_cptb", npr.cptb)):
        np.testing.assert_allclose(d[k][hit], ref, atol=1e-12, rtol=0)
        assert np.isnan(d[k][~hit]).all()
    np.testing.assert_allclose(d["st_dpth"][hit], npr.dpth, atol=1e-12, rtol=0)
    assert np.isnan(d["st_dpth"][~hit]).all()
    np.testing.assert_array_equal(d["st_bids"][hit], npr.bids)
    assert (d["st_bids"][~hit] == -1).all()
