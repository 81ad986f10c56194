"""Autotuners of the unbinding: the direct-summation / particle-mesh
crossover and the FFT convolutions' launch shape.

The reference switches from direct summation to the particle mesh above
1000 members (bound_fortran.f90:1976-1980), a choice made for CPU cores.
On the GPU the balance differs, so ``tune_direct_max`` measures it: for
isolated NFW halos of a ladder of sizes it times the unbinding of a batch
of same-size halos both ways (direct summation forced with a high
threshold, the mesh forced with the reference's) and returns the largest
size at which direct summation is still the cheaper method. The result is
a value for ``Context.set_option(OPT_DIRECT_MAX, ...)``.

``tune_fft_lines`` picks, per mesh size, how many grid lines a workgroup of
the FFT convolutions transforms (4, 8 or 16): more lines read longer row
segments, fewer keep more workgroups resident. The transform of a line does
not depend on the choice, so neither do the results; only the time does.
"""
import numpy as np

from . import _native, _fof_core
from .synth import isolated_halos

__all__ = ["tune_direct_max", "tune_fft_lines", "DEFAULT_LADDER", "FFT_SIZES", "FFT_LINES"]

DEFAULT_LADDER = (1000, 2000, 4000, 8000, 16000, 32000, 65536)


def _time_unbinding(ctx, xs, vs, direct_max, f32, repeats):
    ctx.set_option(_native.OPT_DIRECT_MAX, direct_max)
    ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
    kw = dict(L_Box=512.0, linking_length=0.2, np_halo_min=10, bound_check=True,
              softening=0.1, ng_global=4096.0, _ctx=ctx)
    _fof_core(xs, vs, **kw)  # warm: workspace and Green functions
    best = float("inf")
    for _ in range(repeats):
        _fof_core(xs, vs, **kw)
        best = min(best, ctx.timings()[4] * 1e-3)  # unbinding phase, HIP events
    return best


def tune_direct_max(device=0, f32=True, ladder=DEFAULT_LADDER, halos_per_size=16, repeats=2,
                    seed=77, verbose=False):
    """Returns (direct_max, table) where table lists, per size, the measured
    unbinding time per halo (ms) with direct summation and with the mesh."""
    ctx = _native.Context(device)
    table = []
    best = ladder[0]
    try:
        for n in ladder:
            sizes = np.full(halos_per_size, int(n))
            xs, vs = isolated_halos(seed=seed + int(n), sizes=sizes, interlopers=0.0)
            t_dir = _time_unbinding(ctx, xs, vs, 65536, f32, repeats) / halos_per_size
            t_pm = _time_unbinding(ctx, xs, vs, 1000, f32, repeats) / halos_per_size
            table.append({"n": int(n), "direct_ms": round(t_dir * 1e3, 4),
                          "mesh_ms": round(t_pm * 1e3, 4)})
            if verbose:
                print(f"[tune] n={n}: direct {t_dir * 1e3:.3f} ms/halo, mesh {t_pm * 1e3:.3f} ms/halo")
            if t_dir <= t_pm:
                best = int(n)
    finally:
        ctx.close()
    return best, table


FFT_SIZES = (48, 64, 128, 216, 256)
FFT_LINES = (4, 8, 16)


def _fft_ms(ctx, xs, vs, repeats):
    kw = dict(L_Box=512.0, linking_length=0.2, np_halo_min=10, bound_check=True,
              softening=0.1, ng_global=4096.0, _ctx=ctx)
    _fof_core(xs, vs, **kw)
    best = float("inf")
    for _ in range(repeats):
        _fof_core(xs, vs, **kw)
        best = min(best, ctx.timings()[6])
    return best


def tune_fft_lines(device=0, xs=None, vs=None, repeats=3, min_gain=0.03, seed=13, n_halos=4000,
                   verbose=False):
    """Returns (lines, table): lines maps each mesh size used by the tuning
    set to its fastest lines per workgroup, as values for
    ``Context.set_option(OPT_FFT_LINES, 1000 * n + lines)``; table lists the
    measured FFT-convolution time per call for every (size, lines).

    The tuning set (xs, vs) defaults to n_halos isolated NFW halos of the
    benchmark's generator under another seed, whose sizes reach every mesh
    class. The FFT convolutions are timed with HIP events (OPT_TIME_FFT), the
    best of `repeats` calls. One size is varied at a time with the others at
    8 lines, so a difference in the total is that size's own; a size leaves
    8 lines only for a gain of more than `min_gain` of the total."""
    if xs is None:
        xs, vs = isolated_halos(n_halos=n_halos, seed=seed)
    ctx = _native.Context(device)
    table, lines = [], {}
    try:
        ctx.set_option(_native.OPT_TIME_FFT, 1)
        base = _fft_ms(ctx, xs, vs, repeats)
        st = ctx.stats()
        present = {48: st["n_pm48"] > 0, 64: st["n_pm64"] > 0, 128: st["n_pm128"] > 0,
                   216: st["n_pm256"] > 0, 256: st["n_pm256"] > 0}
        for n in FFT_SIZES:
            if not present[n]:
                continue
            best_ms, best_l = base, 8
            for nl in FFT_LINES:
                ms = base
                if nl != 8:
                    ctx.set_option(_native.OPT_FFT_LINES, 1000 * n + nl)
                    ms = _fft_ms(ctx, xs, vs, repeats)
                table.append({"n": n, "lines": nl, "fft_ms": round(ms, 4)})
                if verbose:
                    print(f"[tune] fft n={n} lines={nl}: {ms:.3f} ms per call")
                if ms < best_ms * (1.0 - min_gain):
                    best_ms, best_l = ms, nl
            ctx.set_option(_native.OPT_FFT_LINES, 1000 * n + 8)
            lines[n] = best_l
    finally:
        ctx.close()
    return lines, table
