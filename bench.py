#!/usr/bin/env python3
"""Benchmark: FoF + unbinding on one MI355X.

Headline (``value``): plain FoF (binning + linking + relabel), 134,217,728
particles, float64 (3, N), L_Box = 512, 50 % in 250,000 NFW clumps (c = 8,
Pareto(0.9) counts in [20, 2e6], r_vir = 0.6 (n/100)^(1/3)), 50 % uniform;
linking_length 0.2, np_halo_min 10, bound_check False. Positions are
generated on the host (seeded), copied to HBM once, and every timed step is
one full fof() call through the C-ABI on device-resident input and outputs.

``unbinding`` (after the headline steps): 20,000 isolated NFW halos, sizes
from dN/dn ~ n^-1.9 on [20, 1e6] (seed 3: 5,803,808 particles), 10 %
interlopers, cube units, softening 0.1, bound_check True; every step is one
fof(bound_check=True) call over the 20,000 regions on device-resident
input. Reported as bound halos/s and particles/s.

Prints one JSON line (see the keys below). The CPU baselines are the oracle
(the C restatement of the reference) timed on bounded samples: FoF on the
same generator at 4,194,304 particles in a 161^3 box (same density) with the
reference's thread rule (2 linking threads for one region); unbinding on the
first halos of the same halo set.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FoF Mparticles/s; bound halos/s batched vs plain at stated err; p99 latency"
HBM_PEAK_GBS = 8000.0
# measured HBM bytes per call of the bbox + binning + linking kernels on the
# 134,217,728-particle workload: the total line of this rocprofv3 summary
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fof_nfw134M_hbm_traffic.md")
# Oracle FoF on the full 134,217,728-particle set against its bounded
# 4,194,304-particle sample, both on the same 8 threads of the build host
# (tests/golden/make_c2_fixture.py, tests/golden/c2_nfw134M.json): 306.4 s
# and 4.6 s, so the full set links at 0.48 times the sample's per-particle
# rate (its clumps reach 2e6 members and denser cores).
CPU_FULL_OVER_SAMPLE = (134_217_728 / 306.4) / (4_194_304 / 4.6)
PHASES = ["bbox", "binning", "linking", "relabel", "unbinding", "catalog"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=134_217_728, help="particles")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=4_194_304, help="CPU baseline sample size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bound", action="store_true", help="skip the unbinding leg")
    ap.add_argument("--bound-halos", type=int, default=20_000)
    ap.add_argument("--bound-steps", type=int, default=3)
    ap.add_argument("--bound-cpu-halos", type=int, default=4_000)
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the plain / f32 / tuned unbinding forms")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-jobs", type=int, default=2000)
    ap.add_argument("--serve-pool", type=int, default=40, help="distinct job inputs, reused")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="Poisson arrival rate, jobs/s (0: 0.9 x the sequential rate)")
    ap.add_argument("--only", choices=["all", "fof", "bound", "serve", "copy"], default="all",
                    help="run one leg only (profiling); the headline needs 'all' or 'fof'")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, msg):
    if not args.quiet:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)


def box_for(n):
    # same mean density (1 particle per unit cell) as 2^27 particles in 512^3
    return 512.0 * (n / 134_217_728) ** (1.0 / 3.0)


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def traffic_from_profile(path=TRAFFIC_PROFILE):
    """Bytes per call from the summary's '**X GB per call**' line, or None."""
    import re
    try:
        with open(path) as f:
            m = re.search(r"\*\*([0-9.]+) GB per call\*\*", f.read())
        return float(m.group(1)) * 1e9 if m else None
    except OSError:
        return None


def cpu_baseline(args):
    """The oracle (C restatement of the reference) on a bounded sample of the
    same generator, with the reference's thread rule for one region (2
    linking threads whatever n_threads is, fof_fortran.f90:369-379) at
    n_threads = the host's cores, and once more linking on every core."""
    from oracle import oracle_fof, oracle_set_link_threads
    from py_fof_amd.synth import clumpy_nfw
    model, cores = host_cpu()
    n = args.cpu_n
    L = box_for(n)
    xp = clumpy_nfw(n=n, L=L, seed=args.seed)
    t0 = time.perf_counter()
    oracle_fof(xp, [n], L, 0.2, 10, n_threads=cores)
    dt = time.perf_counter() - t0
    oracle_set_link_threads(cores)
    try:
        t0 = time.perf_counter()
        oracle_fof(xp, [n], L, 0.2, 10, n_threads=cores)
        dt_all = time.perf_counter() - t0
    finally:
        oracle_set_link_threads(0)
    rate = n / dt / 1e6
    return {
        "value": round(rate, 4), "unit": "Mparticles/s", "cores": 2, "kind": "port",
        "sample": f"oracle FoF on the same generator at {n} particles, L={L:g} (same density), "
                  f"seed {args.seed}, n_threads={cores} (the host's cores), which the "
                  f"reference's rule for one region turns into 2 linking threads; {dt:.2f} s",
        "host": {"cpu": model, "cores_available": cores},
        "all_cores": {"value": round(n / dt_all / 1e6, 4), "unit": "Mparticles/s",
                      "cores": cores, "seconds": round(dt_all, 2),
                      "note": "same sample linked on every core (not the reference's rule)"},
        "full_size_estimate": {
            "value": round(rate * CPU_FULL_OVER_SAMPLE, 4), "unit": "Mparticles/s",
            "factor": round(CPU_FULL_OVER_SAMPLE, 3),
            "note": "sample rate x the full-set / sample per-particle rate ratio of the oracle, "
                    "both measured on 8 threads of the build host (306.4 s for 134,217,728 "
                    "particles, 4.6 s for the 4,194,304-particle sample)"},
    }


def bound_workload(n_halos, seed=3):
    from py_fof_amd.synth import isolated_halos
    xs, vs = isolated_halos(n_halos=n_halos, seed=seed)
    np_halo = np.array([x.shape[1] for x in xs], np.int32)
    xp = np.asfortranarray(np.hstack(xs))
    vp = np.asfortranarray(np.hstack(vs))
    # the reference's default s_15 for a 512 cube with box 300, h0 0.67; it
    # covers every FoF candidate of this set (largest span 16.6 cells)
    s15 = 15.0 * 0.67 * 512.0 / 300.0
    return xp, vp, np_halo, s15


def physical_workload(n_halos, seed=4):
    """The isolated-halo generator in physical units: box 300 Mpc/h,
    n_sim 1024^3, z 0.5, Hubble flow, masses log-normal (0.3 dex) around
    2e9 Msun/h, softening 0.01 Mpc/h (py_fof_amd.synth.physical_halos)."""
    from py_fof_amd.synth import physical_halos
    return physical_halos(n_halos=n_halos, seed=seed)


def bench_physical(args, _native, dmax):
    """Physical-units unbinding: the reference form against the batched f32
    form at the tuned crossover (host inputs through the Python front end)."""
    import py_fof_amd as pf
    log(args, f"physical units: generating {args.bound_halos} halos")
    xs, vs, mp, kw = physical_workload(args.bound_halos)
    n = int(sum(x.shape[1] for x in xs))
    ctx = _native.Context(0)
    res = {}
    try:
        for name, dm, f32 in (("reference_fp64", 1000, False), ("batched_f32_tuned", dmax, True)):
            ctx.set_option(_native.OPT_DIRECT_MAX, dm)
            ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
            ctx.set_option(_native.OPT_RECORD_PHI, 0)
            pf._fof_core(xs, vs, mp_list=mp, _ctx=ctx, **kw)  # warm-up
            t0 = time.perf_counter()
            inh, _, _ = pf._fof_core(xs, vs, mp_list=mp, _ctx=ctx, **kw)
            dt = time.perf_counter() - t0
            st = ctx.stats()
            res[name] = {"bound_halos_per_s": round(st["nvalid"] / dt, 1),
                         "ms_per_call": round(dt * 1e3, 3), "crossover": dm,
                         "unbinding_ms": round(ctx.timings()[4], 3), "in_halo": inh}
        # potential error of the f32 form against FP64 at the same crossover
        phis = []
        for f32 in (False, True):
            ctx.set_option(_native.OPT_DIRECT_MAX, dmax)
            ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
            ctx.set_option(_native.OPT_RECORD_PHI, 1)
            pf._fof_core(xs, vs, mp_list=mp, _ctx=ctx, **kw)
            phis.append(ctx.potential())
    finally:
        ctx.close()
    (p64, i64), (p32, i32) = phis
    rel = np.abs(p32.astype(np.float64) - p64) / np.abs(p64.astype(np.float64))
    diff = int(np.sum(res["reference_fp64"].pop("in_halo") != res["batched_f32_tuned"].pop("in_halo")))
    res.update({
        "workload": f"{args.bound_halos} isolated NFW halos (generator of the cube-unit set, seed 4), "
                    f"{n} particles, box 300 Mpc/h, n_sim 1024^3, z 0.5, include_hubble_flow, "
                    "masses log-normal 0.3 dex around 2e9 Msun/h, softening 0.01 Mpc/h, host inputs",
        "speedup": round(res["reference_fp64"]["ms_per_call"] / res["batched_f32_tuned"]["ms_per_call"], 2),
        "potential_error_f32_vs_fp64": {"members": int(len(p32)),
                                        "same_members": bool(np.array_equal(i64, i32)),
                                        "max_rel": float(f"{rel.max():.3e}"),
                                        "rms_rel": float(f"{np.sqrt(np.mean(rel ** 2)):.3e}")},
        "status_differs_vs_reference_form": {"particles": diff, "fraction": float(f"{diff / n:.3e}")},
    })
    return res


def bench_bound(args, torch, _native, ctx):
    L, b, app = 512.0, 0.2, 0.1
    log(args, f"unbinding: generating {args.bound_halos} isolated halos")
    xp, vp, np_halo, s15 = bound_workload(args.bound_halos)
    n = int(np_halo.sum())
    dev = torch.device("cuda:0")
    x_d = torch.from_numpy(xp.T.copy()).to(dev)
    v_d = torch.from_numpy(vp.T.copy()).to(dev)
    in_h = torch.empty(n, dtype=torch.int32, device=dev)
    hid = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def form(direct_max, f32, batched, record=False):
        ctx.set_option(_native.OPT_DIRECT_MAX, direct_max)
        ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
        ctx.set_option(_native.OPT_BATCHED, int(batched))
        ctx.set_option(_native.OPT_RECORD_PHI, int(record))

    def step():
        err = ctx.fof(x_d, v_d, None, np_halo, L, b, 10, 1, 0, app, 1.0, 0.0, s15, 0, n,
                      256, 512.0, in_h, hid, None, None, _native.MEM_DEVICE)
        if err:
            raise RuntimeError(f"pyfof ierr={err}: {_native.last_error()}")

    same = []

    fft = {"ms": 0.0, "bytes": 0.0, "slots": 0, "calls": 0}

    def timed(steps, time_fft=False):
        ctx.set_option(_native.OPT_TIME_FFT, int(time_fft))
        step()  # warm-up of this form
        first = in_h.clone()
        torch.cuda.synchronize()
        ph = np.zeros(6)
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            tm = ctx.timings()
            ph += np.array(tm[:6])
            if time_fft:
                sf = ctx.stats()
                fft["ms"] += tm[6]
                fft["bytes"] += sf["fft_bytes"]
                fft["slots"] += sf["fft_slots"]
                fft["calls"] += 1
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        same.append(bool(torch.equal(first, in_h)))
        ctx.set_option(_native.OPT_TIME_FFT, 0)
        return dt, ph / steps, ctx.stats()

    # the reference's form: crossover 1000, FP64 direct sums, batched
    form(1000, False, True)
    dt, phases, st = timed(args.bound_steps, time_fft=True)
    ref_in = in_h.clone()
    log(args, f"unbinding (reference form): {dt * 1e3:.1f} ms/call")
    ncand = st["n_ok"] + st["n_too_few"] + st["n_max_iter"]
    out = {
        "workload": f"{args.bound_halos} isolated NFW halos, n ~ n^-1.9 on [20, 1e6], seed 3, "
                    f"{n} particles, 10% interlopers, cube units, b {b}, softening {app}, "
                    f"s_15 {s15:.3f}, bound_check True",
        "bound_halos_per_s": round(st["nvalid"] / dt, 1),
        "candidates_per_s": round(ncand / dt, 1),
        "Mparticles_per_s": round(n / dt / 1e6, 3),
        "ms_per_call": round(dt * 1e3, 3),
        "phases_ms": {k: round(float(v), 3) for k, v in zip(PHASES, phases)},
        "candidates": ncand,
        "bound_halos": st["nvalid"], "status_ok": st["n_ok"], "status_too_few": st["n_too_few"],
        "status_max_iter": st["n_max_iter"], "direct_sum_halos": st["n_direct"],
        "pm_halos": {"48": st["n_pm48"], "64": st["n_pm64"], "128": st["n_pm128"],
                     "256": st["n_pm256"]},
        "pm_passes": st["passes"],
        "form": "reference: direct sum (<= 1000 members) FP64, PM f32, batched",
    }
    if fft["calls"]:
        ach = fft["bytes"] / (fft["ms"] * 1e-3) / 1e9
        out["roofline"] = {
            "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_fft_conv: the FFT convolutions of the mesh potentials, the largest "
                      "share of a call; algorithmic bytes 48 n^3 per slot convolution (forward "
                      "and inverse transforms, three axis passes reading and writing the n^3 "
                      "complex grid), timed with HIP events around each convolution launch",
            "ms_per_call": round(fft["ms"] / fft["calls"], 3),
            "slot_convolutions_per_call": fft["slots"] // fft["calls"],
            "GB_per_call": round(fft["bytes"] / fft["calls"] / 1e9, 3),
        }
    if not args.no_forms:
        from py_fof_amd.tune import tune_direct_max, tune_fft_lines
        # plain per-halo FP64 form (one candidate at a time)
        form(1000, False, False)
        dt_plain, _, st_p = timed(1)
        log(args, f"unbinding (plain FP64 per halo): {dt_plain * 1e3:.1f} ms/call")
        # autotuned crossover and FFT launch shape for the f32 form, then the
        # batched f32 form
        dmax, table = tune_direct_max(f32=True)
        log(args, f"unbinding: tuned direct-summation threshold {dmax}")
        fft_lines, fft_table = tune_fft_lines()
        log(args, f"unbinding: tuned FFT lines per workgroup {fft_lines}")
        form(dmax, True, True)
        for nsz, nl in fft_lines.items():
            ctx.set_option(_native.OPT_FFT_LINES, 1000 * nsz + nl)
        dt_mix, _, st_m = timed(args.bound_steps)
        for nsz in fft_lines:
            ctx.set_option(_native.OPT_FFT_LINES, 1000 * nsz + 8)
        mix_in = in_h.clone()
        log(args, f"unbinding (batched f32, crossover {dmax}): {dt_mix * 1e3:.1f} ms/call")
        # potential error of the f32 form against FP64 at the same crossover
        form(dmax, False, True, record=True)
        step()
        phi64, idx64 = ctx.potential()
        form(dmax, True, True, record=True)
        step()
        phi32, idx32 = ctx.potential()
        form(1000, False, True)
        same_members = bool(np.array_equal(idx64, idx32))
        rel = np.abs(phi32.astype(np.float64) - phi64) / np.abs(phi64.astype(np.float64))
        diff = int((mix_in != ref_in).sum().item())
        out["forms"] = {
            "plain_fp64_per_halo": {"bound_halos_per_s": round(st_p["nvalid"] / dt_plain, 1),
                                    "ms_per_call": round(dt_plain * 1e3, 3),
                                    "crossover": 1000},
            "batched_fp64": {"bound_halos_per_s": out["bound_halos_per_s"],
                             "ms_per_call": out["ms_per_call"], "crossover": 1000},
            "batched_f32_tuned": {"bound_halos_per_s": round(st_m["nvalid"] / dt_mix, 1),
                                  "ms_per_call": round(dt_mix * 1e3, 3), "crossover": dmax,
                                  "fft_lines": {str(k): v for k, v in fft_lines.items()},
                                  "direct_sum_halos": st_m["n_direct"],
                                  "direct_sum_above_1000": st_m["n_direct_large"]},
            "speedup_batched_f32_vs_plain_fp64": round(dt_plain / dt_mix, 2),
            "potential_error_f32_vs_fp64": {
                "members": int(len(phi32)), "same_members": same_members,
                "max_rel": float(f"{rel.max():.3e}"),
                "rms_rel": float(f"{np.sqrt(np.mean(rel ** 2)):.3e}"),
                "note": f"first-pass potential of every candidate member, crossover {dmax}; "
                        "PM members are identical in both forms"},
            "status_differs_vs_reference_form": {
                "particles": diff, "fraction": float(f"{diff / n:.3e}")},
            "tuning": table,
            "fft_lines_tuned": {str(k): v for k, v in fft_lines.items()},
            "fft_lines_tuning": fft_table,
        }
        out["physical_units"] = bench_physical(args, _native, dmax)
    out["deterministic"] = all(same)  # repeated calls of each form gave identical in_halo
    if not args.no_cpu and args.bound_cpu_halos > 0:
        from oracle import oracle_fof_bound
        # bounded sample: the first halos of the set below 20,000 members (the
        # 256^3 two-level halos would take minutes each in the f64 oracle)
        starts = np.concatenate([[0], np.cumsum(np_halo)])
        pick = [i for i in range(len(np_halo)) if np_halo[i] < 20000][:args.bound_cpu_halos]
        cols = np.concatenate([np.arange(starts[i], starts[i + 1]) for i in pick])
        k, nk = len(pick), int(len(cols))
        log(args, f"unbinding cpu baseline on {k} halos ({nk} particles)")
        t0 = time.perf_counter()
        r = oracle_fof_bound(np.asfortranarray(xp[:, cols]), np_halo[pick], L, b, 10,
                             np.asfortranarray(vp[:, cols]), app=app, s15=s15)
        dtc = time.perf_counter() - t0
        nvalid = int(len(np.unique(r["halo_id"][r["halo_id"] > 0])))
        model, cores = host_cpu()
        # the GPU on exactly the same sample (reference form), device-resident
        xs_d = torch.from_numpy(np.ascontiguousarray(xp[:, cols].T)).to(dev)
        vs_d = torch.from_numpy(np.ascontiguousarray(vp[:, cols].T)).to(dev)
        in_s = torch.empty(nk, dtype=torch.int32, device=dev)
        id_s = torch.empty(nk, dtype=torch.int32, device=dev)
        nps = np.ascontiguousarray(np_halo[pick], dtype=np.int32)
        form(1000, False, True)

        def sample_step():
            err = ctx.fof(xs_d, vs_d, None, nps, L, b, 10, 1, 0, app, 1.0, 0.0, s15, 0, nk,
                          256, 512.0, in_s, id_s, None, None, _native.MEM_DEVICE)
            if err:
                raise RuntimeError(f"pyfof ierr={err}: {_native.last_error()}")

        sample_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            sample_step()
        torch.cuda.synchronize()
        dtg = (time.perf_counter() - t0) / 3
        gpu_in = in_s.cpu().numpy()
        ndiff = int(np.sum(gpu_in != r["in_halo"]))
        out["cpu_baseline"] = {
            "value": round(nvalid / dtc, 2), "unit": "bound halos/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle FoF + unbinding on the first {k} halos of the same set with "
                      f"fewer than 20,000 members ({nk} particles, {nvalid} bound halos); "
                      f"FoF linking, direct sums and the removal loops on 1 thread, the FFTs "
                      f"and Green functions on up to {cores}; {dtc:.2f} s",
            "host": {"cpu": model, "cores_available": cores},
            "gpu_same_sample": {"bound_halos_per_s": round(nvalid / dtg, 1),
                                "ms_per_call": round(dtg * 1e3, 3),
                                "speedup_vs_cpu": round(dtc / dtg, 1)},
            "bound_status_vs_oracle": {"particles": ndiff, "of": nk,
                                       "fraction": float(f"{ndiff / max(nk, 1):.3e}")},
        }
    return out


def serve_pool(n_pool, seed=5):
    """Job inputs of the serving stream: periodic sub-boxes of N particles,
    N log-uniform in [1e4, 2e7], the NFW-clump / uniform 50/50 mixture at one
    particle per cube cell (L = N^(1/3)), Gaussian velocities (sigma 0.5),
    30 % of them (rounded) with bound_check (s_15 from ng_global 2048)."""
    from py_fof_amd.synth import clumpy_nfw
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(1e4), np.log(2e7), n_pool)).astype(np.int64)
    bound = np.zeros(n_pool, bool)
    bound[rng.choice(n_pool, int(round(0.3 * n_pool)), replace=False)] = True
    jobs = []
    for k, (n, b) in enumerate(zip(ns, bound)):
        L = float(n) ** (1.0 / 3.0)
        xp = clumpy_nfw(n=int(n), L=L, seed=seed * 1000 + k)
        vp = np.asfortranarray(rng.normal(0.0, 0.5, (3, int(n))))
        kw = dict(L_Box=L, linking_length=0.2, np_halo_min=10, bound_check=bool(b),
                  ng_global=2048.0)
        jobs.append((xp, vp, kw))
    return jobs


def _pct(v, q):
    return round(float(np.percentile(np.asarray(v), q)), 3)


def bench_serve(args):
    import threading
    from py_fof_amd.serve import Server
    log(args, f"serving: generating a pool of {args.serve_pool} jobs")
    pool = serve_pool(args.serve_pool)
    sizes = [j[0].shape[1] for j in pool]
    with Server(slots=1, hbm_budget=200 << 30) as srv:
        # sequential plain calls: one job at a time, warm workspace
        for xp, vp, kw in pool[:2]:
            srv.submit(xp, vp, **kw).result()
        seq = []
        t0 = time.perf_counter()
        for xp, vp, kw in pool:
            t = time.perf_counter()
            srv.submit(xp, vp, **kw).result()
            seq.append(time.perf_counter() - t)
        seq_wall = time.perf_counter() - t0
    seq_rate = len(pool) / seq_wall
    log(args, f"serving: sequential {seq_rate:.2f} jobs/s over the pool")
    # capacity: every job of 3 passes over the pool submitted at once
    with Server(slots=args.serve_slots, hbm_budget=200 << 30) as srv:
        for xp, vp, kw in pool[:args.serve_slots]:
            srv.submit(xp, vp, **kw).result()
        t0 = time.perf_counter()
        hs = [srv.submit(xp, vp, **kw) for _ in range(3) for xp, vp, kw in pool]
        for h in hs:
            h.result()
        cap_rate = len(hs) / (time.perf_counter() - t0)
    log(args, f"serving: capacity with {args.serve_slots} slots {cap_rate:.2f} jobs/s")
    rate = args.serve_rate if args.serve_rate > 0 else 0.9 * seq_rate
    rng = np.random.default_rng(55)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, args.serve_jobs))
    pick = rng.integers(0, len(pool), args.serve_jobs)  # which pool input each job carries
    lat = [None] * args.serve_jobs
    handles = [None] * args.serve_jobs
    with Server(slots=args.serve_slots, hbm_budget=200 << 30) as srv:
        for xp, vp, kw in pool[:args.serve_slots]:
            srv.submit(xp, vp, **kw).result()
        done = threading.Event()

        def collect():
            for i in range(args.serve_jobs):
                while handles[i] is None:
                    time.sleep(0.0005)
                h, t_arr = handles[i]
                h.result()
                lat[i] = time.perf_counter() - t_arr
            done.set()

        th = threading.Thread(target=collect, daemon=True)
        th.start()
        t0 = time.perf_counter()
        last_log = t0
        for i in range(args.serve_jobs):
            wait = t0 + arrivals[i] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            xp, vp, kw = pool[pick[i]]
            t_arr = t0 + arrivals[i]
            handles[i] = (srv.submit(xp, vp, **kw), t_arr)
            if time.perf_counter() - last_log > 20.0:
                last_log = time.perf_counter()
                log(args, f"serving: {i + 1}/{args.serve_jobs} jobs submitted")
        done.wait()
        wall = time.perf_counter() - t0
    lat_ms = [x * 1e3 for x in lat]
    nb = sum(1 for j in pool if j[2]["bound_check"])
    nbj = int(sum(1 for k in pick if pool[k][2]["bound_check"]))
    return {
        "workload": f"{args.serve_jobs} jobs, Poisson arrivals at {rate:.2f} jobs/s, inputs "
                    f"from a pool of {len(pool)} periodic sub-boxes (N log-uniform in [1e4, 2e7], "
                    f"NFW clumps 50% + uniform, 1 particle per cell, sigma_v 0.5; {nb} of "
                    f"{len(pool)} with bound_check), each job a uniformly drawn pool input "
                    f"({nbj} jobs with bound_check), host inputs, {args.serve_slots} slots",
        "arrival_rate_jobs_per_s": round(rate, 3),
        "jobs_per_s": round(args.serve_jobs / wall, 3),
        "p50_ms": _pct(lat_ms, 50),
        "p99_ms": _pct(lat_ms, 99),
        "mean_ms": round(float(np.mean(lat_ms)), 3),
        "capacity_jobs_per_s": round(cap_rate, 3),
        "sequential_jobs_per_s": round(seq_rate, 3),
        "sequential_p50_ms": _pct([x * 1e3 for x in seq], 50),
        "sequential_p99_ms": _pct([x * 1e3 for x in seq], 99),
        "pool_particles": int(sum(sizes)),
        "latency": "arrival to result, including queueing; sequential: one call at a time",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from py_fof_amd import _native
    from py_fof_amd.synth import clumpy_nfw

    if args.only == "bound":
        ctx = _native.Context(0)
        try:
            res = bench_bound(args, torch, _native, ctx)
        finally:
            ctx.close()
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "unbinding": res}), flush=True)
        return
    if args.only == "copy":
        # calibration for HBM counters: a device copy of 1 GiB (1 GiB read,
        # 1 GiB written per launch), 3 launches
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 3
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                          "ms_per_copy": round(dt * 1e3, 3),
                          "GB_per_s": round(2 * (1 << 30) / dt / 1e9, 1)}), flush=True)
        return
    if args.only == "serve":
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "serving": bench_serve(args)}),
              flush=True)
        return

    n = args.n
    L = box_for(n)
    log(args, f"generating {n} particles, L={L:g}")
    t = time.time()
    xp = clumpy_nfw(n=n, L=L, seed=args.seed)
    log(args, f"generated in {time.time() - t:.1f} s")
    dev = torch.device("cuda:0")
    x_d = torch.from_numpy(xp.T).to(dev)  # (N, 3) row-major == (3, N) Fortran order
    del xp
    in_h = torch.empty(n, dtype=torch.int32, device=dev)
    hid = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx = _native.Context(0)
    np_halo = np.array([n], np.int32)

    def step():
        err = ctx.fof(x_d, None, None, np_halo, L, 0.2, 10, 0, 0, 0.1, 1.0, 0.0, 17.152, 0, n,
                      256, 512.0, in_h, hid, None, None, _native.MEM_DEVICE)
        if err:
            raise RuntimeError(f"pyfof ierr={err}: {_native.last_error()}")

    for i in range(args.warmup):
        step()
        log(args, f"warmup {i} done: phases ms {[round(v, 2) for v in ctx.timings()]}")
    ref_in = in_h.clone()
    torch.cuda.synchronize()
    phases = np.zeros(6)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        phases += np.array(ctx.timings()[:6])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(args, f"timed {args.steps} steps in {t1 - t0:.3f} s")
    ms = (t1 - t0) * 1e3 / args.steps
    phases /= args.steps
    st = ctx.stats()
    same = bool(torch.equal(ref_in, in_h))
    ngroups = st.get("nhalo", 0)
    # compulsory HBM traffic of binning + linking: N*64 B + cells*12 B
    alg_bytes = n * 64 + st["ncell"] * 12
    t_bl = (phases[0] + phases[1] + phases[2]) * 1e-3
    achieved = alg_bytes / t_bl / 1e9
    out = {
        "metric": METRIC,
        "value": round(n / (ms * 1e-3) / 1e6, 3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32",
        "data": "synthetic",
        "config": {
            "workload": "plain FoF, 134,217,728 particles (NFW clumps 50% + uniform), "
                        "L_Box 512, b 0.2, np_halo_min 10, bound_check False"
                        if n == 134_217_728 else f"plain FoF, {n} particles, L_Box {L:g}",
            "particles": n, "L_Box": L, "linking_length": 0.2, "np_halo_min": 10,
            "parallelism": "single GPU", "input": "device-resident f64 (3,N)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile() if n == 134_217_728 else None,
            "kernel": "binning + linking (k_bbox..k_dense_cross), algorithmic bytes N*64 + cells*12; "
                      "traffic: measured HBM bytes of the same kernels per call",
        },
        "phases_ms": {k: round(float(v), 3) for k, v in zip(
            PHASES, phases)},
        "groups": ngroups,
        "cells": st["ncell"],
        "dense_cells": st["ndense"],
        "in_halo": int(in_h.sum().item()),
        "deterministic": same,
    }
    if not args.no_cpu:
        log(args, "cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args)
    if not args.no_bound and args.only == "all":
        del x_d, in_h, hid
        torch.cuda.empty_cache()
        out["unbinding"] = bench_bound(args, torch, _native, ctx)
    ctx.close()
    if not args.no_serve and args.only == "all":
        out["serving"] = bench_serve(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
