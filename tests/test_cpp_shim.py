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
from _common import assert_track_parity
from _find_common import best_hit_reference, explain_best_hit
from oracle.parity import describe, fit_reference, flagged_parity

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
    o, o64, chk = fit_reference(3.8, rl, xi, ev, 2, 1)
    assert_track_parity("shim fit", (res[:6], res[6:27], res[27]), (o[0], o[1], o[3]), (o64[0], o64[1], o64[3]), chk)


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
    geo = mk.geometry(1)
    ev = mk.generate_find_event(geo, mk.default_toy_params(seed=seed, simulate_material=0, pt_min=0.5, pt_max=10.0),
                                n, noise_per_cm2=3.0)
    assert ev["n_seeds"] == ns
    cpu, chk = best_hit_reference(ev, mk.barrel_layer_descs(geo), 3, chi2_cut=15.0, bfield=3.8, n_threads=8)
    # both C++ paths against the oracle: hit lists explained seed by seed, parameters per track (the shim
    # returns parameters only)
    for name, hot, par in (("shim layer loop", hot_a, par_a), ("shim fused", hot_b, par_b)):
        gpu = dict(hot_idx=hot, par=par, err=None, chi2=None, n_found=cpu["n_found"], chg=cpu["chg"])
        n_diff, counts, bad = explain_best_hit(hot, cpu["hot_idx"], cpu["margins"], 3)
        for t, l in bad[:3]:
            print(name, "seed", t, "layer", l, "gpu", hot[:, t].tolist(), "oracle", cpu["hot_idx"][:, t].tolist(),
                  "fused", hot_b[:, t].tolist(), "margins", cpu["margins"][:, l, t].tolist())
        assert not bad, (name, n_diff, counts, bad[:10])
        same = np.all(hot == cpu["hot_idx"], axis=0)
        r = flagged_parity((par[:, same], None, None), (cpu["par"][:, same], cpu["err"][:, same], cpu["chi2"][same]),
                           (cpu["par"][:, same], cpu["err"][:, same], cpu["chi2"][same]),
                           {"spread": chk["spread"][:, same], "alt": [tuple(x[..., same] for x in a) for a in chk["alt"]],
                            "tie_all": True})
        print(name, n_diff, counts, describe(r))
        assert r["ok"], (name, r["beyond"])
