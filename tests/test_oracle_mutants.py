"""The oracle pins have teeth: each planted mistake must make some pin fail.

Every mutant is a one-place textual change to a copy of the oracle source (a
dropped term, a flipped sign, a transposed operand, a wrong index), compiled
into a temporary library and run through every pin of tests/pins.py.  A
mutant that passes all pins would mean that part of the restatement is not
pinned; the test fails and names it.
"""
import concurrent.futures
import os
import shutil
import subprocess
import tempfile

import pytest

import oracle
import pins

ORACLE_DIR = os.path.dirname(oracle.__file__)

MUTANTS = {
    # propagation: helix step and its Jacobian (PropagationMPlex.icc)
    "jac03_sign": ("J(0, 3) = k * (cosPorT", "J(0, 3) = -k * (cosPorT"),
    "jac14_drop_term": ("+ sinPorT * cosa + cosPorT * sina - sinPorT) * pt;", "+ sinPorT * cosa - sinPorT) * pt;"),
    "jac44_sign": ("J(4, 4) = T(1) + dadphi;", "J(4, 4) = T(1) - dadphi;"),
    "jac25_power": ("J(2, 5) = -k * alpha * pt * sinPorT * sinPorT;", "J(2, 5) = -k * alpha * pt * sinPorT;"),
    "dDdphi_sign": ("dDdphi += k * (x * (pysa", "dDdphi -= k * (x * (pysa"),
    "sincos4_drop_x4": ("c = T(1) - T(0.5f) * x2 + T(0.04166667f) * x2 * x2;", "c = T(1) - T(0.5f) * x2;"),
    "fail_no_negative": ("if (oodotp > T(5.0f) || oodotp < T(0)) {", "if (oodotp > T(5.0f)) {"),
    "helixz_j05": ("J(0, 5) = deltaZ * ipt * pxcaMpysa * icos2T;", "J(0, 5) = deltaZ * ipt * pxcaMpysa;"),
    "helixz_j43_sign": ("J(4, 3) = tanT * deltaZ * kinv;", "J(4, 3) = -tanT * deltaZ * kinv;"),
    # similarity
    "similarity_transposed_blocked": ("const T* a = J + (j * 6 + k) * NN;", "const T* a = J + (k * 6 + j) * NN;"),
    # material (PropagationMPlex.cc:716-767)
    "highland_log_term": ("(T(1) + T(0.038f) * std::log(radL))", "(T(1))"),
    "eloss_sign": ("const T dP = propSign[n] * dEdx / beta;", "const T dP = -propSign[n] * dEdx / beta;"),
    "ms_theta_only": ("outErr[sym(4, 4) * NN + n] += thetaMSC2;", ""),
    # Kalman update (KalmanUtilsMPlex.cc)
    "rotation_sign": ("const T a00 = -(my + py) / (2 * r);", "const T a00 = (my + py) / (2 * r);"),
    "gain_index": ("K[i][1] = C(i, 0) * h[1] + C(i, 1) * h[4] + C(i, 2) * h[7];",
                   "K[i][1] = C(i, 0) * h[1] + C(i, 1) * h[4] + C(i, 2) * h[6];"),
    "chi2_cross_term": ("2 * (s1 * l1 * l0)", "(s1 * l1 * l0)"),
    "khc_drop_z": ("KH[i][0] * C(0, j) + KH[i][1] * C(1, j) + KH[i][2] * C(2, j);",
                   "KH[i][0] * C(0, j) + KH[i][1] * C(1, j);"),
    "endcap_gain_swap": ("K[i][1] = C(i, 0) * s1 + C(i, 1) * s2;", "K[i][1] = C(i, 1) * s1 + C(i, 0) * s2;"),
    "inverse_offdiag_sign": ("a1 *= -s;", "a1 *= s;"),
}


def _build(name, old, new, workdir):
    d = os.path.join(workdir, name)
    os.makedirs(d)
    for f in ("oracle.cc", "oracle.h", "oracle_core.h"):
        shutil.copy(os.path.join(ORACLE_DIR, f), d)
    p = os.path.join(d, "oracle_core.h")
    src = open(p).read()
    assert src.count(old) == 1, "%s: pattern matches %d times" % (name, src.count(old))
    open(p, "w").write(src.replace(old, new))
    so = os.path.join(d, "liborc_mut.so")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread", "-DORC_NN=8", "-o", so,
                        os.path.join(d, "oracle.cc")], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return so


@pytest.fixture(scope="module")
def mutant_libs():
    work = tempfile.mkdtemp(prefix="orc_mut_")
    try:
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
            futs = {k: ex.submit(_build, k, a, b, work) for k, (a, b) in MUTANTS.items()}
            yield {k: f.result() for k, f in futs.items()}
    finally:
        shutil.rmtree(work, ignore_errors=True)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(mutant_libs, name):
    so = mutant_libs[name]

    def make(bfield, rl, xi):
        return oracle.Oracle(bfield, rl, xi, lib_path=so)

    caught = []
    for pin in pins.ALL_PINS:
        try:
            pin(make)
        except Exception as e:  # AssertionError, or NaN-driven numpy errors
            caught.append("%s: %s" % (pin.__name__, type(e).__name__))
    assert caught, "mutant %s passes every pin" % name
