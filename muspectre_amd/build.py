"""Build step of the package: compiles libmugpu.so (HIP, gfx950), the
pybind11 module _muSpectre and the CPU oracle used by the tests.

Run as ``python muspectre_amd/build.py`` (make decides per object file what
is out of date; ``--force`` rebuilds everything) or through
__graft_entry__.build(), which always rebuilds everything.
"""
import glob
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_hip(force=False, jobs=8):
    csrc = os.path.join(HERE, "csrc")
    if force:
        _run(["make", "clean"], cwd=csrc)
    _run(["make", "-j{}".format(jobs)], cwd=csrc)


def module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_muSpectre" + suffix)


def build_module():
    import pybind11
    host = os.path.join(HERE, "csrc", "host")
    srcs = sorted(glob.glob(os.path.join(host, "*.cc")))
    out = module_path()
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I" + os.path.join(ROOT, "include"), "-I" + host,
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           *srcs, "-L" + os.path.join(HERE, "lib"), "-lmugpu",
           "-Wl,-rpath,$ORIGIN/lib", "-o", out]
    _run(cmd)
    return out


def build_oracle(force=False):
    _run(["make", "-C", os.path.join(ROOT, "oracle")] + (["-B"] if force else []))


def build_all(force=False):
    build_hip(force=force)
    build_module()
    build_oracle(force=force)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
