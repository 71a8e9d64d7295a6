"""The C++ interface (include/mkfit_amd.hpp): it compiles against the C ABI, and on the GPU a
MkFitter-style caller (tests/cpp/fit_shim_demo.cc) reproduces the oracle's fit of the same event, and
a MkBuilder-style best-hit layer loop over the per-layer MkFinder steps (tests/cpp/find_shim_demo.cc)
reproduces the one-call FindTracksBestHit and the oracle's."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import mkfit_amd as mk
import oracle
from _common import par_rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "fit_shim_demo")


@pytest.mark.parametrize("name", ["fit_shim_demo", "find_shim_demo"])
def test_shim_compiles_and_links(name):
    from mkfit_amd import _build
    if not os.path.exists(os.path.join(ROOT, "mkfit_amd", "libmkfit_amd.so")):
        pytest.skip("library not built")
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "demo")
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I",
                            os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-o", out,
                            os.path.join(ROOT, "tests", "cpp", name + ".cc"), "-L", _build.PKG,
                            "-lmkfit_amd", "-lmkfit_toymc", "-L", "/opt/rocm/lib", "-lamdhip64"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["batched", "plain"])
def test_shim_fit_matches_oracle(gpu_device, form, tmp_path):
    n, seed = 20000, 99
    out = str(tmp_path / "fit.bin")
    cmd = [DEMO, out, str(n), str(seed)] + (["plain"] if form == "plain" else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(out, np.float32).reshape(28, n)
    geo = mk.geometry(1)
    rl, xi = mk.material_grid(geo)
    ev = mk.generate(geo, mk.default_toy_params(seed=seed), n, geo.n_barrel, n_threads=4)
    O = oracle.Oracle(bfield=3.8, rl_grid=rl, xi_grid=xi)
    par, err, chg, chi2 = O.fit(ev, pflags=2, fit_flags=1, n_threads=8)
    pr = par_rel(res[:6], par)
    assert np.mean(pr.max(axis=0) <= 1e-5) > 0.99
    c2 = np.abs(res[27] - chi2) / np.maximum(1.0, np.abs(chi2))
    assert np.percentile(c2, 95) < 1e-4


@pytest.mark.gpu
def test_shim_layer_loop_matches_find_best_hit(gpu_device, tmp_path):
    n, seed = 3000, 55
    out = str(tmp_path / "find.bin")
    r = subprocess.run([os.path.join(ROOT, "tests", "cpp", "find_shim_demo"), out, str(n), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.int32)
    ns, L = int(raw[0]), int(raw[1])
    body = raw[2:]
    blk = L * ns + 6 * ns
    hot_a, hot_b = body[:L * ns].reshape(L, ns), body[blk:blk + L * ns].reshape(L, ns)
    par_a = body[L * ns:blk].view(np.float32).reshape(6, ns)
    par_b = body[blk + L * ns:2 * blk].view(np.float32).reshape(6, ns)
    same = np.all(hot_a == hot_b, axis=0)
    assert same.mean() >= 0.995, same.mean()
    pr = par_rel(par_b[:, same], par_a[:, same])
    assert np.percentile(pr, 99) < 1e-5, np.percentile(pr, 99)
    geo = mk.geometry(1)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=seed, simulate_material=0, pt_min=0.5, pt_max=10.0),
                                n, noise_per_cm2=3.0)
    assert ev["n_seeds"] == ns
    cpu = oracle.find_best_hit(ev, mk.barrel_layer_descs(geo), first_layer=3, chi2_cut=15.0, bfield=3.8,
                               n_threads=8)
    assert np.all(hot_b == cpu["hot_idx"], axis=0).mean() >= 0.99
