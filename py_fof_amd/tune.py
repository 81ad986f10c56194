"""Autotuner of the direct-summation / particle-mesh crossover.

The reference switches from direct summation to the particle mesh above
1000 members (bound_fortran.f90:1976-1980), a choice made for CPU cores.
On the GPU the balance differs, so ``tune_direct_max`` measures it: for
isolated NFW halos of a ladder of sizes it times the unbinding of a batch
of same-size halos both ways (direct summation forced with a high
threshold, the mesh forced with the reference's) and returns the largest
size at which direct summation is still the cheaper method. The result is
a value for ``Context.set_option(OPT_DIRECT_MAX, ...)``.
"""
import numpy as np

from . import _native, _fof_core
from .synth import isolated_halos

__all__ = ["tune_direct_max", "DEFAULT_LADDER"]

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
