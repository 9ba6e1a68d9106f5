"""Autotuner of the compile-time-length FFT passes.

Every pass of the operator (last-axis r2c with the p update, the previous
iteration's x update and K:p,
middle-axis c2c, first-axis c2c with the projection, last-axis c2r with the
p.Ap partials, first-axis c2c of the engine transforms) exists in two
variants for the power-of-two lengths: 16 complex values per thread, or 8
with twice the threads on the same tile (bit 0 of the variant).  The
last-axis passes and the first-axis projection pass also come with half the
rows or k-points per workgroup (bit 1): half the LDS, more workgroups per CU
(and, on the first axis, shorter HBM runs).  The first-axis projection pass
also exists with an occupancy hint on its register allocation (bit 2),
which changes its instruction schedule.  Which one
is fastest depends on the pass, the length, the dimension and the
precision.  This script times every variant of every pass on the GPU
(mugpu_fft_tune_pass, zeroed scratch buffers of a grid large enough to fill
the device) and writes the table the library is compiled with
(csrc/fft_tuned.inc).  A variant replaces the default only when it is
faster by more than the margin, so the table does not flip on timing noise;
the table is part of the source, so results are the same from run to run.

    python -m muspectre_amd.autotune --out muspectre_amd/csrc/fft_tuned.inc \\
        --summary tuning.md
"""
import argparse
import ctypes
import os

from . import _capi

PASSES = {0: "r2c + K:p", 1: "middle c2c", 2: "axis-0 c2c + projection", 3: "c2r + p.Ap",
          4: "axis-0 c2c (engine)"}
LENGTHS = [16, 32, 64, 128, 256, 512]


def grid(pass_, n, dim):
    """A grid with the tuned axis of length n: the cube n^3 for n = 128 and
    256 (the large single-RVE grids), otherwise other axes of 256 points, large
    enough to fill the device (short axes stand for many small RVEs served in
    one batched launch; 512^3 behaves like 256 x 256 x 512)."""
    if dim == 3:
        m = n if n in (128, 256) else 256
        return {0: [m, m, n], 3: [m, m, n], 1: [m, n, m], 2: [n, m, m], 4: [n, m, m]}[pass_]
    m = 512  # the longest outer axis with compiled passes
    return {0: [m, n], 3: [m, n], 2: [n, m], 4: [n, m]}[pass_]


def time_pass(lib, res, prec, pass_, variant, iso, reps):
    r = (ctypes.c_int * len(res))(*res)
    ms = ctypes.c_double()
    _capi.check(lib.mugpu_fft_tune_pass(len(res), r, prec, pass_, variant, iso, reps,
                                        ctypes.byref(ms)))
    return ms.value


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True, help="path of the generated table")
    ap.add_argument("--summary", default=None, help="markdown summary of the timings")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--margin", type=float, default=0.03,
                    help="relative gain a variant needs to replace the default")
    ap.add_argument("--min-gain-ms", type=float, default=0.002,
                    help="absolute gain a variant needs to replace the default")
    args = ap.parse_args()
    lib = _capi.load()
    _capi.check(lib.mugpu_init(0))
    rows, entries = [], []
    for dim in (3, 2):
        for prec, pname in ((_capi.FP64, "fp64"), (_capi.FP32, "fp32")):
            for pass_ in (0, 1, 2, 3, 4):
                if pass_ == 1 and dim == 2:
                    continue
                for iso in ((0, 1) if pass_ == 0 else (0,)):
                    for n in LENGTHS:
                        res = grid(pass_, n, dim)
                        t = {}
                        for v in (range(8) if pass_ == 2 else (0, 1, 2, 3) if pass_ in (0, 3)
                                  else (0, 1)):
                            try:
                                t[v] = time_pass(lib, res, prec, pass_, v, iso, args.reps)
                            except _capi.MugpuError:
                                pass  # no such variant at this length
                        best = 0
                        for v, tv in t.items():
                            if (v and tv < (1.0 - args.margin) * t[0] and t[0] - tv > args.min_gain_ms
                                    and tv < t[best]):
                                best = v
                        rows.append((dim, pname, PASSES[pass_], iso, n, res, t, best))
                        print("dim {} {} {:26s} iso {} n {:4d} grid {}: {} ms -> {}".format(
                            dim, pname, PASSES[pass_], iso, n, res,
                            " / ".join("{:.3f}".format(t[v]) if v in t else "-" for v in range(8)),
                            best), flush=True)
                        if best:
                            entries.append((pass_, n, dim, prec, iso, best))
    lines = ["// Variants of the compile-time-length FFT passes: {pass, axis length, dim,",
             "// precision (0 FP64, 1 FP32), isotropic tangent (r2c pass only), variant}.",
             "// Written by muspectre_amd/autotune.py; entries absent here take variant 0.",
             "static const FastTuned kFastTuned[] = {",
             "    {-1, 0, 0, 0, 0, 0},"]
    lines += ["    {%d, %d, %d, %d, %d, %d}," % e for e in entries]
    lines.append("};")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if args.summary:
        with open(args.summary, "w") as f:
            f.write("# FFT pass variants chosen by the autotuner\n\nCommand (on one MI355X):\n\n"
                    "    python3 -u -m muspectre_amd.autotune --out muspectre_amd/csrc/fft_tuned.inc "
                    "--summary profiles/fft_autotune.md\n\n"
                    "Average of {} launches per variant (HIP events around each launch, after "
                    "a 1 GiB copy that evicts the last-level cache), zeroed scratch buffers; a "
                    "variant with 8 values per thread replaces the default (16, or 12 for lengths "
                    "with a factor 3) only when it is faster by more than {:g} % and {:g} us. The "
                    "table is compiled into libmugpu.so (muspectre_amd/csrc/fft_tuned.inc).\n\n".format(
                        args.reps, 100 * args.margin, 1000 * args.min_gain_ms))
            f.write("Variants: 0 = 16 (12) values per thread, 1 = 8 values per thread, "
                    "2 and 3 = the same with half the rows (last-axis passes) or k-points "
                    "(first-axis projection pass) per workgroup, 4-7 = variants 0-3 of the "
                    "first-axis projection pass with the occupancy hint; - = no such variant "
                    "at this length.\n\n")
            f.write("| dim | precision | pass | isotropic tangent | n | grid | ms, variant 0 | "
                    "1 | 2 | 3 | 4 | 5 | 6 | 7 | chosen |\n"
                    "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|\n")
            for dim, pname, pn, iso, n, res, t, best in rows:
                f.write("| {} | {} | {} | {} | {} | {} | {} | {} |\n".format(
                    dim, pname, pn, iso, n, "x".join(map(str, res)),
                    " | ".join("{:.3f}".format(t[v]) if v in t else "-" for v in range(8)), best))
    print("wrote {} ({} entries)".format(os.path.abspath(args.out), len(entries)))


if __name__ == "__main__":
    main()
