"""Build recipe for the native libraries (no cmake; direct hipcc / g++ lines).

  mkfit_amd/libmkfit_amd.so    HIP kernels + C ABI (include/mkfit_amd.h), gfx950, and the event-file
                               reader/writer (include/mkfit_event_file.h, host code)
  mkfit_amd/libmkfit_toymc.so  toy Monte Carlo generator (include/mkfit_toymc.h), host C++
  oracle/liborc_avx512.so      CPU oracle, NN=16 (test infrastructure)
  oracle/liborc_avx2.so        CPU oracle, NN=8  (test infrastructure)
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mkfit_amd")
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed FP32 (v_pk_*_f32) is kept out of generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_hip():
    out = os.path.join(PKG, "libmkfit_amd.so")
    srcs = [os.path.join(CSRC, "mkf_fit.hip"), os.path.join(CSRC, "mkf_find.hip"),
            os.path.join(CSRC, "mkf_ce.hip"), os.path.join(CSRC, "mkf_finder_layer.hip"),
            os.path.join(CSRC, "mkf_adjoint.hip"), os.path.join(CSRC, "mkf_bkfit.hip"),
            os.path.join(CSRC, "event_file.cc")]
    _run([HIPCC] + HIP_FLAGS + ["-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_toymc():
    out = os.path.join(PKG, "libmkfit_toymc.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", out + ".tmp",
          os.path.join(CSRC, "toymc.cc")])
    os.replace(out + ".tmp", out)
    return out


def build_shim_demo():
    """tests/cpp/{fit,find}_shim_demo: the C++ interface (include/mkfit_amd.hpp) driven like MkFitter and
    like MkBuilder's best-hit layer loop."""
    outs = []
    for name in ("fit_shim_demo", "find_shim_demo"):
        out = os.path.join(ROOT, "tests", "cpp", name)
        _run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
              "-I", "/opt/rocm/include", "-o", out + ".tmp", os.path.join(ROOT, "tests", "cpp", name + ".cc"),
              "-L", PKG, "-lmkfit_amd", "-lmkfit_toymc", "-L", "/opt/rocm/lib", "-lamdhip64",
              "-Wl,-rpath," + PKG, "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-j4", "-C", odir, "-B", "all"])


def build_all():
    build_oracle()
    build_toymc()
    build_hip()
    build_shim_demo()


if __name__ == "__main__":
    build_all()
    print("built")
