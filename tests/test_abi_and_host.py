"""CPU checks of the product boundary: the C-ABI library loads and exports
every symbol declared in include/mugpu.h, the pybind11 module imports and
exposes the reference's Python surface, and without a GPU every call that
needs the device fails loudly (no CPU fallback)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mugpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mugpu_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "muspectre_amd", "lib", "libmugpu.so"))
    names = _declared()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_capi_signatures_cover_header():
    from muspectre_amd import _capi
    declared = set(_declared())
    bound = {s[0] for s in _capi.SIGNATURES}
    assert declared == bound


def test_capi_argument_lists_match_header():
    """Every ctypes binding passes as many arguments as its declaration takes,
    a pointer where the declaration takes one and a scalar of the same width
    where it does not."""
    import ctypes
    from muspectre_amd import _capi
    src = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "mugpu.h")).read(), flags=re.S)
    decl = {}
    for m in re.finditer(r"\b(mugpu_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        a = m.group(2).strip()
        decl[m.group(1)] = [] if a in ("", "void") else [x.strip() for x in a.split(",")]
    handles = {"mugpu_cell", "mugpu_fft", "mugpu_proj", "mugpu_server"}
    widths = {"int": 4, "int64_t": 8, "size_t": 8, "double": 8, "float": 4}
    for name, _, args in _capi.SIGNATURES:
        params = decl[name]
        assert len(params) == len(args), name
        for c, t in zip(params, args):
            words = c.replace("*", " * ").split()
            base = [w for w in words if w not in ("const", "*")][0]
            if "*" in words or base in handles:
                assert t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer), (name, c)
            else:
                assert ctypes.sizeof(t) == widths[base] and t not in (ctypes.c_void_p,), (name, c)
                assert (t in (ctypes.c_double, ctypes.c_float)) == (base in ("double", "float")), (name, c)


def test_module_surface():
    import muspectre_amd as mu
    for name in ("Cell", "Formulation", "material", "solvers", "fft", "get_domain_ccoord",
                 "get_domain_index", "get_hermitian_sizes", "OptimizeResult"):
        assert hasattr(mu, name)
    for d in ("2d", "3d"):
        for m in ("MaterialLinearElastic1", "MaterialLinearElastic2", "MaterialLinearElastic3",
                  "MaterialLinearElastic4", "MaterialHyperElastoPlastic1"):
            assert hasattr(mu.material, m + "_" + d)
        for p in ("ProjectionSmallStrain", "ProjectionFiniteStrain", "ProjectionFiniteStrainFast"):
            assert hasattr(mu._muSpectre.fft, p + "_" + d)
        assert hasattr(mu._muSpectre.fft, "FFTW_" + d)
    assert hasattr(mu.solvers, "SolverCG") and hasattr(mu.solvers, "newton_cg")
    assert hasattr(mu.solvers, "de_geus")


def test_grid_helpers():
    import muspectre_amd as mu
    assert mu.get_hermitian_sizes([5, 7, 9]) == [5, 7, 5]
    assert mu.get_hermitian_sizes([6, 4]) == [6, 3]
    assert mu.get_domain_ccoord([3, 4], 5) == [1, 1]
    assert mu.get_domain_index([3, 4], [1, 1]) == 5
    assert mu.get_domain_index([3, 4, 5], [2, 3, 4]) == 59


def test_no_gpu_fails_loudly():
    """With no device the product raises instead of computing on the CPU."""
    import torch  # noqa: F401  (only to ask whether a device exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import muspectre_amd as mu
    with pytest.raises(RuntimeError):
        mu.Cell([5, 5], [1., 1.], mu.Formulation.small_strain)
    with pytest.raises(RuntimeError):
        mu.fft.FFT([5, 5], 4).initialise()


def test_error_text_is_per_thread():
    """mugpu_last_error() gives the calling thread's last message: a failure
    on another thread does not change it.  (Argument checks of
    mugpu_server_footprint run before any device call.)"""
    import threading
    lib = ctypes.CDLL(os.path.join(ROOT, "muspectre_amd", "lib", "libmugpu.so"))
    lib.mugpu_last_error.restype = ctypes.c_char_p
    size = ctypes.c_size_t()

    def fail(n):
        res = (ctypes.c_int * 3)(n, 32, 32)
        assert lib.mugpu_server_footprint(3, res, ctypes.byref(size)) != 0
        return lib.mugpu_last_error().decode()

    mine = fail(17)
    assert "got 17" in mine
    seen = []
    t = threading.Thread(target=lambda: seen.append((lib.mugpu_last_error().decode(), fail(19))))
    t.start()
    t.join()
    before, theirs = seen[0]
    assert before == "" and "got 19" in theirs
    assert lib.mugpu_last_error().decode() == mine


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N != 1 exits with status 2 and says why, before
    it touches a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_product_never_touches_the_oracle():
    """Only the build script names the oracle (it compiles the checker beside
    the product); no runtime file of the package imports or loads it."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "muspectre_amd")):
        for f in files:
            if f == "build.py":
                continue
            if f.endswith((".py", ".cc", ".hh", ".hip", ".hpp")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in text.lower(), f
