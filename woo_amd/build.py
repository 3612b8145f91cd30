"""Build recipe for the HIP engine library ``woo_amd/libwoo_amd.so`` (gfx950).

Every kernel is compiled with ``-ffp-contract=off`` (no fused multiply-add, so
FP64 results are reproducible bit for bit) and with packed FP32 instructions
disabled. Run ``python -m woo_amd.build`` or ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwoo_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WOO_AMD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]

SOURCES = ["wdg_sort.hip", "wdg_collide.hip", "wdg_contact.hip", "wdg_integrate.hip", "wdg_pack.hip", "wdg_graph.hip", "wdg_mpnn.hip",
           "wdg_capi.hip"]
HEADERS = ["wdg_internal.hpp", "wdg_math.hpp", "wdg_math32.hpp", "wdg_physics.hpp"]


def _newest(paths):
    return max(os.path.getmtime(p) for p in paths)


def _compile(src, force):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in HEADERS] + [
        os.path.join(ROOT, "include", "woo_amd.h")]
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= _newest(deps):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    os.replace(obj + ".tmp", obj)
    return obj


HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["woo_host.cpp", "woo_py.cpp"]
HOST_HEADERS = ["woo_host.hpp"]


def host_module_path():
    import sysconfig

    return os.path.join(HERE, "_woo" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_host(force=False, verbose=True):
    """The host classes (Scene, DemField, engines, functors) as the Python module
    woo_amd._woo: plain C++ (g++) against pybind11, linked to libwoo_amd.so."""
    import sysconfig

    import pybind11

    out = host_module_path()
    deps = [os.path.join(HOST, f) for f in HOST_SOURCES + HOST_HEADERS] + [os.path.join(ROOT, "include", "woo_amd.h"),
                                                                         LIB]
    if not force and os.path.exists(out) and os.path.getmtime(out) >= _newest(deps):
        return out
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"], "-I", os.path.join(ROOT, "include")]
    cmd += [os.path.join(HOST, f) for f in HOST_SOURCES]
    cmd += ["-L", HERE, "-lwoo_amd", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host module build failed:\n{r.stderr[-6000:]}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")
    return out


PROBE_SRC = os.path.join(HERE, "probe", "history_probe.hip")
PROBE_LIB = os.path.join(HERE, "libwdg_probe.so")


def build_probe(force=False, verbose=True):
    """The history-access probe bench.py --history-probe uses (a measurement, not the engine)."""
    if not force and os.path.exists(PROBE_LIB) and os.path.getmtime(PROBE_LIB) >= os.path.getmtime(PROBE_SRC):
        return PROBE_LIB
    cmd = [HIPCC] + FLAGS + ["-shared", PROBE_SRC, "-o", PROBE_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for the history probe:\n{r.stderr[-6000:]}")
    os.replace(PROBE_LIB + ".tmp", PROBE_LIB)
    if verbose:
        print(f"built {PROBE_LIB}")
    return PROBE_LIB


def build(force=False, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < _newest(objs):
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
        os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    build_host(force=force, verbose=verbose)
    build_probe(force=force, verbose=verbose)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
