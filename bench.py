"""Benchmark: acoustic forward shot on one MI355X.

Workload (default): 2001 (z) x 4001 (x) interior + 50-cell C-PML on every
side (4101 x 2101 cells stored), dx = dz = 10 m, dt = 0.8 ms, nt = 6000,
velocity = seeded Gaussian-smoothed noise (sigma = 20 cells) mapped to
1500-4500 m/s, constant density, 20 Hz Ricker at interior (z=2, x=2000), 4001
receivers at interior z=2, radius-8 staggered stencil (the reference's C1..C8).

One "step" of this benchmark is one whole shot (nt time steps) with the model,
wavelet and geometry already resident in device memory.  value = Mcell-updates/s
over all stored cells (boundary layers included).

Launch shapes: the library's autotuner (sfd_acoustic_plan_autotune /
sfd_acoustic_batch_autotune) times every built shape (60 steps for the
single shot, 20 for the batch) before the timed shots and keeps the fastest (--no-tune: the library's defaults); the
chosen shape is reported as launch_shape.

Roofline: the library runs each time step as two launches (velocities, then
pressure; acoustic_tiled.h), so the roofline's unit is one time step:
achieved = the fused step's algorithmic bytes (28 B per cell with the
workload's uniform density, plus the C-PML strips' memory variables) / the
device time of one step (both launches),
measured with HIP events on the library's stream over the timed shots;
traffic = both launches' FETCH_SIZE + WRITE_SIZE per step, scaled by a
256 MiB copy (profiles/acoustic_step_cpml_2101x4101.md).

FP64 errors: every leg states the relative L2 distance of its result from
the FP64 oracle on a stated sample (the headline: its gather on the full grid
over the first 800 steps), next to the FP32 oracle's own distance.  The
streaming ceiling (hbm_ceiling) is measured by hand-written read / write /
copy kernels in the same run; roofline.ceiling_frac = achieved / copy rate.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "Mcell-updates/s & % HBM roofline; batched fwd+adjoint shots/s vs plain fwd"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_step(shot, plan_pml_w=None):
    """Compulsory HBM bytes of one time step: P, Vx, Vz read and written
    (24 B), the per-cell coefficients read (kf and bf, 8 B; kf only, 4 B,
    when the density is uniform: the library then passes bf as one number),
    and for C-PML the two memory variables of each strip cell read and
    written (16 B)."""
    nx, nz = shot.vp.shape
    uniform = shot.den is None or bool(np.all(np.asarray(shot.den) == np.asarray(shot.den).flat[0]))
    b = (28.0 if uniform else 32.0) * nx * nz
    if shot.boundary == "cpml":
        w = shot.width + 1
        xrows = min(2 * w, nx)
        zcols = min(2 * w, nz)
        b += 16.0 * (xrows * nz + zcols * nx)
    return b


def cpu_baseline(shot, steps):
    import oracle
    cores = min(16, len(os.sched_getaffinity(0)))
    den = np.full_like(shot.vp, 2400.0)
    extra = {}
    if shot.boundary == "cpml":
        extra = dict(boundary="cpml", pml_r=shot.pml_r, pml_f0=shot.pml_f0, pml_vmax=float(shot.vp.max()))
    t0 = time.perf_counter()
    oracle.acoustic(shot.vp, den, shot.wavelet, shot.src, shot.rec, dx=shot.dx, dz=shot.dz, dt=shot.dt,
                    radius=shot.radius, sponge=shot.width, nt=steps, nthreads=cores, **extra)
    dt = time.perf_counter() - t0
    nx, nz = shot.vp.shape
    return {"value": round(nx * nz * steps / dt / 1e6, 2), "unit": "Mcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"first {steps} of {len(shot.wavelet)} time steps of the same shot, FP32 oracle "
                      f"(reference arithmetic), OpenMP over x rows, {dt:.1f} s"}


def c1_prefix_error(shot, steps):
    """FP64 error of the headline shot on its full grid over the first
    `steps` steps: the gather of a plan run that far against the FP64 and
    FP32 oracles."""
    import oracle
    import seismic_code_github_amd as sfd
    cores = min(16, len(os.sched_getaffinity(0)))
    short = sfd.AcousticShot(**{**shot.__dict__, "wavelet": shot.wavelet[:steps], "nt": None, "_keep": []})
    rec, _ = sfd.acoustic_forward(short)
    den = np.full_like(shot.vp, 2400.0)
    kw = dict(dx=shot.dx, dz=shot.dz, dt=shot.dt, radius=shot.radius, sponge=shot.width, nt=steps, nthreads=cores,
              boundary="cpml", pml_r=shot.pml_r, pml_f0=shot.pml_f0, pml_vmax=float(shot.vp.max()))
    o64 = oracle.acoustic(shot.vp, den, shot.wavelet, shot.src, shot.rec, precision="f64", **kw)["record"]
    o32 = oracle.acoustic(shot.vp, den, shot.wavelet, shot.src, shot.rec, **kw)["record"]
    nrm = np.linalg.norm(o64)
    return {"fp64_rel_l2": float(np.linalg.norm(rec - o64) / nrm),
            "fp32_oracle_vs_fp64_rel_l2": float(np.linalg.norm(o32 - o64) / nrm),
            "error_sample": f"gather of the same shot on its full {shot.vp.shape[0]}x{shot.vp.shape[1]} grid over "
                            f"the first {steps} of {len(shot.wavelet)} steps"}


STREAM_MIN_ROWS = 64  # one-launch shapes of this many rows or more are the streaming step


def shape_dict(kind, a, b, us):
    if kind == 0:
        d = {"kernel": "tiled two-launch step", "tile_rows": a, "waves": b}
    elif kind == 2 and a >= STREAM_MIN_ROWS:
        d = {"kernel": "streaming one-launch step", "chunk_rows": a, "waves": b}
    elif kind == 2:
        d = {"kernel": "one-launch tiled step", "tile_rows": a, "waves": b}
    else:
        d = {"kernel": "fused march kernel", "waves": a, "chunk_rows": b}
    d["autotuned_us_per_step"] = None if us is None else round(us, 2)
    return d


def kernel_stem(kind, a, batch):
    """Name stem of the step kernel(s) a launch shape runs."""
    if kind == 0:
        stem = "acoustic_tiled_velocity"
    elif kind == 2:
        stem = "acoustic_stream_step" if a >= STREAM_MIN_ROWS else "acoustic_fused_step"
    else:
        return "acoustic_batch_step_kernel" if batch else "acoustic_step_kernel"
    return stem + ("_batch" if batch else "")


def load_traffic(name, key, stem):
    """A counter profile's HBM bytes, only when the profiled kernels are the
    ones the chosen launch shape runs (else None: the profile is stale)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if stem and not any((stem if stem.endswith(">") or "<" in stem else stem + "<") in k
                        for k in t.get("kernels", [])):
        return None
    return t.get(key)


def batched_leg(nt, no_tune=False, no_cpu=False, plain=True, launches=False):
    """64 shots on uneven aperture windows of a 1001 x 6001 model (C-PML 50,
    nt=3000 per shot): one step per time step for all shots of a group (the
    library's default groups, sized to the MALL, one after another), replayed
    from a hipGraph, against the plain path (shot after shot, one step per
    shot and time step, each at its own grid class's tuned shape).  Launch
    shape from the autotuner.  Timed over `nt` steps (default all 3000).
    FP64 error: the gather of the median-width window against the FP64
    oracle over the same steps."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    _, shots = W.windowed_shots(nshot=64, nt=nt)
    cells = sum(s.vp.size for s in shots)
    bytes_step = sum(algorithmic_bytes_per_step(s) for s in shots)
    b = sfd.AcousticBatch(shots)
    # 60 trial steps: every shot group (the batch runs its shots in groups
    # sized to the MALL) starts its trial from cold caches.
    tuned_us = None if no_tune else b.autotune(60)
    kind, ta, tb = b.shape()
    ngroups = b.groups()
    log(f"batched leg: shape {shape_dict(kind, ta, tb, tuned_us)}, {ngroups} shot groups")
    if plain and not no_tune:
        # The plain path runs every shot at its own grid class's shape: tune
        # each class once on one of its shots.
        seen = set()
        for s in shots:
            cls = tuple(int(round(np.log2(n))) for n in s.vp.shape)
            if cls not in seen:
                seen.add(cls)
                with sfd.AcousticPlan(s) as pl:
                    pl.autotune(20)
    how = sfd.BATCH_LAUNCHES if launches else sfd.BATCH_GRAPH
    b.run(how)
    g_ms = min(b.run(how) for _ in range(2))
    log(f"batched leg: {'graph' if how == sfd.BATCH_GRAPH else 'launches'} {g_ms:.1f} ms for {nt} steps")
    order = np.argsort([s.vp.shape[0] for s in shots])
    mid = int(order[len(shots) // 2])
    rec_mid = b.fetch(mid)
    p_ms = b.run(sfd.BATCH_PER_SHOT) if plain else float("nan")
    b.close()
    scale = 3000.0 / nt
    out = {"workload": "64 shots, windows of 1000-2500 columns of a 1001x6001 model, C-PML 50, R=8, nt=3000",
           "cells": int(cells), "timed_steps": nt, "shot_groups": ngroups,
           "graph_Mcell_per_s": round(cells * nt / g_ms / 1e3, 1),
           "plain_Mcell_per_s": round(cells * nt / p_ms / 1e3, 1),
           "graph_shots_per_s": round(64 / (g_ms * scale / 1e3), 3),
           "plain_shots_per_s": round(64 / (p_ms * scale / 1e3), 3),
           "speedup": round(p_ms / g_ms, 2),
           "bytes_per_step": bytes_step,
           "avg_step_us": round(g_ms * 1e3 / nt, 2),
           # uniform density: 28 algorithmic bytes per cell-update + the C-PML strips
           "graph_frac_of_8TBs": round(bytes_step * nt / (g_ms / 1e3) / 8e12, 4),
           "launch_shape": shape_dict(kind, ta, tb, tuned_us)}
    out["traffic"] = load_traffic("acoustic_batch_c3_traffic.json", "hbm_bytes_per_step", kernel_stem(kind, ta, True))
    if not no_cpu:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        sh = shots[mid]
        kw = dict(dx=sh.dx, dz=sh.dz, dt=sh.dt, radius=sh.radius, sponge=sh.width, nt=nt, nthreads=cores,
                  boundary="cpml", pml_r=sh.pml_r, pml_f0=sh.pml_f0, pml_vmax=float(sh.vp.max()))
        den = np.full_like(sh.vp, 2400.0)
        o64 = oracle.acoustic(sh.vp, den, sh.wavelet, sh.src, sh.rec, precision="f64", **kw)["record"]
        o32 = oracle.acoustic(sh.vp, den, sh.wavelet, sh.src, sh.rec, **kw)["record"]
        nrm = np.linalg.norm(o64)
        out["fp64_rel_l2"] = float(np.linalg.norm(rec_mid - o64) / nrm)
        out["fp32_oracle_vs_fp64_rel_l2"] = float(np.linalg.norm(o32 - o64) / nrm)
        out["error_sample"] = f"gather of window {mid} ({sh.vp.shape[0]}x{sh.vp.shape[1]} cells), all {nt} steps"
    return out


def c0_leg(no_cpu, no_tune=False):
    """301 x 301 interior + 40-cell sponge (381 x 381), 3-layer
    velocity, R = 4 (8th order), dt = 1 ms, nt = 2000, 25 Hz Ricker, 301
    receivers at z = 2: GPU Mcell-updates/s in fast and reference arithmetic
    beside the reference CPU path (the FP32 oracle in the reference's
    operation order) on 1 core and on all host cores, with the FP64 error of
    the gathers.  Cache-resident and launch-bound, so no roofline fraction."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    shot = W.layered_301(radius=4)
    nx, nz = shot.vp.shape
    nt = len(shot.wavelet)
    out = {"workload": "acoustic 301x301 interior + 40-cell sponge (381x381), 3 layers, R=4 (8th order), "
                       "dt=1 ms, nt=2000, 25 Hz Ricker, 301 receivers", "cells": nx * nz, "nt": nt}
    recs = {}
    for arith in ("fast", "reference"):
        sh = W.layered_301(radius=4, arith=arith)
        with sfd.AcousticPlan(sh) as pl:
            us = None if no_tune else pl.autotune(100)
            shape = pl.shape()
            pl.run()
            ms = min(pl.run() for _ in range(3))
            recs[arith] = pl.fetch()[0]
        out[arith] = {"ms_per_shot": round(ms, 3), "Mcell_per_s": round(nx * nz * nt / ms / 1e3, 1),
                      "us_per_step": round(ms * 1e3 / nt, 2), "launch_shape": shape_dict(*shape, us)}
    if not no_cpu:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        den = np.full_like(shot.vp, 2400.0)
        kw = dict(dx=shot.dx, dz=shot.dz, dt=shot.dt, radius=4, sponge=shot.width)
        o64 = oracle.acoustic(shot.vp, den, shot.wavelet, shot.src, shot.rec, precision="f64", nthreads=cores,
                              **kw)["record"]
        base = {}
        for th in (1, cores):
            t0 = time.perf_counter()
            o32 = oracle.acoustic(shot.vp, den, shot.wavelet, shot.src, shot.rec, nthreads=th, **kw)["record"]
            base[th] = time.perf_counter() - t0
        nrm = np.linalg.norm(o64)
        out["fast"]["fp64_rel_l2"] = float(np.linalg.norm(recs["fast"] - o64) / nrm)
        out["reference"]["fp64_rel_l2"] = float(np.linalg.norm(recs["reference"] - o64) / nrm)
        out["reference"]["bitwise_equal_to_cpu_reference_path"] = bool(np.array_equal(recs["reference"], o32))
        out["fp32_cpu_vs_fp64_rel_l2"] = float(np.linalg.norm(o32 - o64) / nrm)
        out["cpu_baseline"] = {"value": round(nx * nz * nt / base[1] / 1e6, 2), "unit": "Mcell-updates/s",
                               "cores": 1, "kind": "port",
                               "value_all_cores": round(nx * nz * nt / base[cores] / 1e6, 2),
                               "all_cores": cores,
                               "sample": f"the whole shot ({nt} steps), FP32 oracle in the reference's operation "
                                         f"order (bitwise the reference program's arithmetic), {base[1]:.2f} s "
                                         f"on 1 core, {base[cores]:.2f} s on {cores}"}
        out["gpu_vs_cpu_1core"] = round(out["fast"]["Mcell_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def visco_leg(nt_err, nt, no_cpu, nbatch=256):
    """Visco-acoustic 1001 x 2001 + 40-cell sponge, L = 3 (one shot, nt = 4000
    recorded + the reference's 60-step delay): Mcell-updates/s over all stored
    cells, the algorithmic bytes of the step kernel, the error against the
    FP64 oracle on a bounded prefix of the same shot (the first `nt_err`
    recorded samples) and the FP32 oracle's rate on `cpu_steps` steps."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    shot = W.visco_1001x2001(nt=nt)
    nx, nz = shot.vp.shape
    L = shot.nmech
    nsteps = shot.ref_delay() + shot.nt
    with sfd.ViscoPlan(shot) as pl:
        tuned_us = pl.autotune(20)
        vshape = pl.shape()
        pl.run()
        ms = min(pl.run() for _ in range(2))
        rec, _ = pl.fetch()
    assert np.isfinite(rec).all()
    sp = shot.sponge
    strip_cells = (sp[0] + sp[1] + 2) * nz + (sp[2] + sp[3] + 2) * nx
    # Per cell: sigma, vx, vz read + written (24 B), r_1..r_L read + written
    # (8 B each), b, ku, kr read (12 B); per sponge-strip cell the x or z parts
    # of sigma and r_l read + written (8 B each).
    bytes_step = (36.0 + 8.0 * L) * nx * nz + 8.0 * (1 + L) * strip_cells
    out = {"workload": f"visco-acoustic RSG, R=5, L={L}, 1001x2001 interior + 40-cell cosine sponge "
                       f"({nx}x{nz} stored), dt=1 ms, nt=4000 (+{shot.ref_delay()} delay steps), 20 Hz",
           "cells": nx * nz, "steps": nsteps, "ms_per_shot": round(ms, 2),
           "Mcell_per_s": round(nx * nz * nsteps / ms / 1e3, 1),
           "bytes_per_cell_update": 36 + 8 * L,
           "achieved_GBs": round(bytes_step * nsteps / (ms / 1e3) / 1e9, 1),
           "frac_of_8TBs": round(bytes_step * nsteps / (ms / 1e3) / 8e12, 4),
           "launch_shape": shape_dict(*vshape, tuned_us)}
    # The same shot with vs > 0 (vs = vp / a seeded smooth ratio in 1.6..2.2,
    # qs = 0.8 qp): the visco-elastic kernels, two launches per step.
    vs = (shot.vp / W.smooth_random_model(nx, nz, 20.0, 1.6, 2.2, 7)).astype(np.float32)
    eshot = W.visco_1001x2001(nt=nt, vs=vs, qs=(0.8 * shot.qp).astype(np.float32))
    with sfd.ViscoPlan(eshot) as pl:
        pl.run()
        ems = min(pl.run() for _ in range(2))
        erec, _ = pl.fetch()
    assert np.isfinite(erec).all()
    xcells = (sp[0] + sp[1] + 2) * nz
    zcells = (sp[2] + sp[3] + 2) * nx
    # Per cell: velocity launch reads 3 stresses, vx, vz, b and writes vx, vz
    # (32 B); stress launch reads vx, vz, 3 stresses, 3L memory variables and
    # 4 moduli, writes 3 stresses and 3L memory variables (48 + 24 L B).  Per
    # strip cell and direction: 2 velocity parts and 3 + 3L stress / memory
    # parts read and written.
    ebytes = (80.0 + 24.0 * L) * nx * nz + 8.0 * (5 + 3 * L) * (xcells + zcells)
    out["visco_elastic"] = {"workload": f"same shot, vs = vp / (1.6..2.2), qs = 0.8 qp, L={L}",
                            "ms_per_shot": round(ems, 2), "Mcell_per_s": round(nx * nz * nsteps / ems / 1e3, 1),
                            "us_per_step": round(ems * 1e3 / nsteps, 2),
                            "bytes_per_cell_update": 80 + 24 * L,
                            "achieved_GBs": round(ebytes * nsteps / (ems / 1e3) / 1e9, 1),
                            "frac_of_8TBs": round(ebytes * nsteps / (ems / 1e3) / 8e12, 4)}
    # The Q_forward card's survey batched: nbatch shots (visco-elastic, L = 1,
    # six window columns + 21-cell sponges each), one launch per half-step for
    # all of them, against the same shots one plan at a time (a sample of 8).
    if nbatch > 0:
        shots = W.visco_card_survey(nshot=nbatch)
        s0 = shots[0]
        bnx, bnz = s0.vp.shape
        bsteps = s0.ref_delay() + s0.nt
        with sfd.ViscoBatch(shots) as b:
            vgroups = b.groups()
            b.run()
            bms = min(b.run() for _ in range(2))
            brec = b.fetch(0)[0]
        assert np.isfinite(brec).all()
        one = []
        for sh in shots[:8]:
            with sfd.ViscoPlan(sh) as pl:
                pl.run()
                one.append(pl.run())
        sms = float(np.mean(one)) * nbatch
        sp = s0.sponge
        xc = min(sp[0] + sp[1] + 2, bnx) * bnz
        zc = min(sp[2] + sp[3] + 2, bnz) * bnx
        Lb = s0.nmech
        bb = ((80.0 + 24.0 * Lb) * bnx * bnz + 8.0 * (5 + 3 * Lb) * (xc + zc)) * nbatch
        out["batch"] = {"workload": f"{nbatch} shots of the Q_forward card's survey (isoq_2d_modelling_simple.par: "
                                    f"401x401 two-layer synthetic model, offsets 0-50 m, 200 m sponges), each "
                                    f"{bnx}x{bnz}, visco-elastic, L={Lb}, nt={s0.nt} (+{s0.ref_delay()} delay steps), "
                                    f"10 Hz",
                        "shots": nbatch, "shot_groups": vgroups, "ms_per_batch": round(bms, 2),
                        "shots_per_s": round(nbatch / (bms / 1e3), 1),
                        "plain_shots_per_s": round(nbatch / (sms / 1e3), 1), "speedup": round(sms / bms, 2),
                        "Mcell_per_s": round(nbatch * bnx * bnz * bsteps / bms / 1e3, 1),
                        "us_per_step": round(bms * 1e3 / bsteps, 2),
                        "bytes_per_step": bb, "achieved_GBs": round(bb * bsteps / (bms / 1e3) / 1e9, 1),
                        "frac_of_8TBs": round(bb * bsteps / (bms / 1e3) / 8e12, 4),
                        "plain_sample": "8 shots one plan at a time (tiled step, two launches per step), scaled"}
        if not no_cpu:
            import oracle
            cores = min(16, len(os.sched_getaffinity(0)))
            kw = dict(src=s0.src, rec=s0.rec, nsteps=bsteps, dx=s0.dx, dz=s0.dz, dt=s0.dt, f0=s0.f0, vs=s0.vs,
                      qs=s0.qs, radius=s0.radius, nmech=Lb, sponge=s0.sponge, rec_delay=s0.ref_delay(),
                      nthreads=cores)
            o64 = oracle.visco(s0.vp, s0.rho, s0.qp, precision="f64", **kw)["rec_zz"]
            o32 = oracle.visco(s0.vp, s0.rho, s0.qp, precision="f32", **kw)["rec_zz"]
            nrm = np.linalg.norm(o64)
            out["batch"]["fp64_rel_l2"] = float(np.linalg.norm(brec - o64) / nrm)
            out["batch"]["fp32_oracle_vs_fp64_rel_l2"] = float(np.linalg.norm(o32 - o64) / nrm)
            out["batch"]["error_sample"] = f"strzz gather of shot 1, all {s0.nt} samples"
        out["batch"]["traffic"] = load_traffic(f"visco_batch_{nbatch}x{bnx}x{bnz}_traffic.json", "hbm_bytes_per_step",
                                               "visco_el_velocity_batch" if s0.elastic() else "visco_ac_velocity_batch")
    if nt == 4000 and vshape[0] == 0:
        out["traffic"] = load_traffic("visco_step_L3_2081x1081_traffic.json", "hbm_bytes_per_launch",
                                      f"visco_ac_velocity_kernel<{shot.radius}, {vshape[1]}, {vshape[2]}>")
    if not no_cpu:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        short = W.visco_1001x2001(nt=nt_err)
        rs, _ = sfd.visco_forward(short)
        args = dict(src=short.src, rec=short.rec, nsteps=short.ref_delay() + nt_err, dt=short.dt, f0=short.f0,
                    nmech=L, sponge=short.sponge, rec_delay=short.ref_delay(), nthreads=cores)
        rho = np.full_like(short.vp, 2400.0)
        o64 = oracle.visco(short.vp, rho, short.qp, precision="f64", **args)["rec_zz"]
        t0 = time.perf_counter()
        o32 = oracle.visco(short.vp, rho, short.qp, precision="f32", **args)["rec_zz"]
        dt32 = time.perf_counter() - t0
        nrm = np.linalg.norm(o64)
        out["fp64_rel_l2"] = float(np.linalg.norm(rs - o64) / nrm)
        out["fp32_oracle_vs_fp64_rel_l2"] = float(np.linalg.norm(o32 - o64) / nrm)
        out["error_sample"] = f"first {nt_err} recorded samples ({args['nsteps']} steps) of the same shot"
        out["cpu_baseline"] = {"value": round(nx * nz * args["nsteps"] / dt32 / 1e6, 2), "unit": "Mcell-updates/s",
                               "cores": cores, "kind": "port",
                               "sample": f"{args['nsteps']} steps of the same shot, FP32 oracle (split fields as "
                                         f"Q_forward stores them), OpenMP over x rows, {dt32:.1f} s"}
    return out


def elastic_leg(nt_err, nt, no_cpu, nbatch=32):
    """Elastic free surface on the reference program's own shot grid
    (iso_2d_modelling_sin_test.par: 903 x 603 padded, nt = 3001 recorded + 48
    delay steps): Mcell-updates/s over all stored cells, the algorithmic bytes
    of the two step kernels, the error against the FP64 oracle on the first
    `nt_err` recorded samples, and the FP32 oracle's rate on the same prefix."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    shot = W.elastic_par_903x603(nt=nt)
    nx, nz = shot.vp.shape
    nsteps = shot.ref_delay() + shot.nt
    with sfd.ElasticPlan(shot) as pl:
        tuned_us = pl.autotune(60)
        eshape = pl.shape()
        pl.run()
        ms = min(pl.run() for _ in range(3))
        rvx, rvz, _ = pl.fetch()
    assert np.isfinite(rvx).all() and np.isfinite(rvz).all()
    sp = shot.sponge
    strip_cells = (sp[0] + sp[1] + 2) * nz + (nx - sp[0] - sp[1] - 2) * (sp[2] + 1)
    # Per cell and step: velocity kernel reads tau_xx, tau_zz, tau_xz, vx, vz,
    # busg2e, busg1e and writes vx, vz (36 B); stress kernel reads vx, vz,
    # tau_xx, tau_zz, tau_xz, lambmusg, lambsg, musge and writes the three
    # stresses (44 B).  Per sponge-strip cell the x and z parts of all five
    # fields are read and written too (80 B).
    bytes_step = 80.0 * nx * nz + 80.0 * strip_cells
    out = {"workload": f"elastic SSG free surface, 10th order, iso_2d_modelling_sin_test.par shot grid "
                       f"{nx}x{nz} (501x401 model + 201-cell cosine sponges), dt=1 ms, nt={nt} "
                       f"(+{shot.ref_delay()} delay steps), 25 Hz, rgau=50",
           "cells": nx * nz, "strip_cells": strip_cells, "steps": nsteps, "ms_per_shot": round(ms, 2),
           "Mcell_per_s": round(nx * nz * nsteps / ms / 1e3, 1),
           "us_per_step": round(ms * 1e3 / nsteps, 2),
           "bytes_per_step": bytes_step,
           "achieved_GBs": round(bytes_step * nsteps / (ms / 1e3) / 1e9, 1),
           "frac_of_8TBs": round(bytes_step * nsteps / (ms / 1e3) / 8e12, 4),
           "launch_shape": {"kernel": "tiled two-launch step", "tile_cols": eshape[0], "waves": eshape[1],
                            "autotuned_us_per_step": round(tuned_us, 2)}}
    tpath = os.path.join(ROOT, "profiles", "elastic_step_903x603_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            out["traffic"] = json.load(f).get("hbm_bytes_per_step")
    # The card's survey batched: shots 1..nbatch (the window follows the
    # source), one launch per half-step over every shot's tiles.
    brec = None
    if nbatch > 1:
        shots = W.elastic_par_survey(nshot=nbatch, nt=nt)
        with sfd.ElasticBatch(shots) as b:
            btuned = b.autotune(20)
            bshape = b.shape()
            bgroups = b.groups()
            b.run()
            bms = min(b.run() for _ in range(2))
            brec = b.fetch(0)[:2]
        assert np.isfinite(brec[0]).all() and np.isfinite(brec[1]).all()
        bbytes = bytes_step * nbatch
        out["batch"] = {"workload": f"{nbatch} shots of the card's survey (shots 1..{nbatch}, dsx = 100 m), "
                                    f"each on the {nx}x{nz} shot grid",
                        "shots": nbatch, "shot_groups": bgroups, "ms_per_batch": round(bms, 2),
                        "shots_per_s": round(nbatch / (bms / 1e3), 2),
                        "Mcell_per_s": round(nbatch * nx * nz * nsteps / bms / 1e3, 1),
                        "us_per_step": round(bms * 1e3 / nsteps, 2),
                        "bytes_per_step": bbytes,
                        "achieved_GBs": round(bbytes * nsteps / (bms / 1e3) / 1e9, 1),
                        "frac_of_8TBs": round(bbytes * nsteps / (bms / 1e3) / 8e12, 4),
                        "speedup_vs_single_shot": round(nbatch * ms / bms, 2),
                        "launch_shape": {"kernel": "batched two-launch step", "tile_cols": bshape[0],
                                         "waves": bshape[1], "autotuned_us_per_step": round(btuned, 2)}}
        bpath = os.path.join(ROOT, "profiles", f"elastic_batch_{nbatch}x903x603_traffic.json")
        if os.path.exists(bpath):
            with open(bpath) as f:
                out["batch"]["traffic"] = json.load(f).get("hbm_bytes_per_step")
    if not no_cpu:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        short = W.elastic_par_903x603(nt=nt_err)
        svx, svz, _ = sfd.elastic_forward(short)
        args = dict(src=short.src, rec=short.rec, nsteps=short.ref_delay() + nt_err, delay=short.ref_delay(),
                    dx=short.dx, dz=short.dz, dt=short.dt, f0=short.f0, sponge=short.sponge, rgau=short.src_radius,
                    nthreads=cores)
        o64 = oracle.elastic(short.vp, short.vs, short.rho, short.itopo, precision="f64", **args)
        t0 = time.perf_counter()
        o32 = oracle.elastic(short.vp, short.vs, short.rho, short.itopo, precision="f32", **args)
        dt32 = time.perf_counter() - t0
        for q, k in enumerate(("vx", "vz")):
            got = (svx, svz)[q]
            ref = o64["rec_" + k]
            nrm = np.linalg.norm(ref)
            out[f"fp64_rel_l2_{k}"] = float(np.linalg.norm(got - ref) / nrm)
            out[f"fp32_oracle_vs_fp64_rel_l2_{k}"] = float(np.linalg.norm(o32["rec_" + k] - ref) / nrm)
            if brec is not None:
                # the batch's shot 1 is this shot: its first nt_err samples
                out["batch"][f"fp64_rel_l2_{k}"] = float(np.linalg.norm(brec[q][:, :nt_err] - ref) / nrm)
        if brec is not None:
            out["batch"]["error_sample"] = f"shot 1 of the batch, first {nt_err} recorded samples"
        out["error_sample"] = f"first {nt_err} recorded samples ({args['nsteps']} steps) of the same shot"
        out["cpu_baseline"] = {"value": round(nx * nz * args["nsteps"] / dt32 / 1e6, 2), "unit": "Mcell-updates/s",
                               "cores": cores, "kind": "port",
                               "sample": f"{args['nsteps']} steps of the same shot, FP32 oracle (split fields as "
                                         f"the reference stores them), OpenMP over x columns, {dt32:.1f} s"}
    return out


def vti_bytes_per_step(nx, nz, pml):
    """Algorithmic HBM bytes of one step of one VTI shot.  Interior cell:
    read p, q (current), p, q (old), vz, eps, delta, write p, q = 36 B.  Left /
    right strip cell: no old p, q read (28 B) plus the split parts: U1, Ut,
    U3, Q1, Qt two levels and U2, Q2 one level read, seven written (76 B).
    Top / bottom strip cell: no old p read (32 B) plus U1, Ut, U3 two levels,
    U2 one level read, four written (44 B)."""
    nl, nr, nu, nd = pml
    xs = (nl + nr) * nz
    zs = (nx - nl - nr) * (nu + nd)
    return 36 * (nx * nz - xs - zs) + (28 + 76) * xs + (32 + 44) * zs


def vti_leg(nshot, nt, nt_err, no_cpu, no_tune=False, tiling=None):
    """VTI qP-acoustic survey of the reference card
    (vti_qa_2d_modelling_new.par): nshot shots of the 698 x 348 padded grid,
    nt recorded samples + 140 delay steps, all shots advanced together by one
    launch per step.  Parity: the first nt_err recorded samples of one shot
    against the FP64 oracle; CPU baseline: the FP32 oracle on that prefix."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    V, E, D, kw = W.vti_card_survey(nshot=nshot, nt=nt)
    _, nx, nz = V.shape
    with sfd.VtiPlan(V, E, D, **kw) as pl:
        tuned_us = None
        if tiling:
            pl.set_tiling(*tiling)
        elif not no_tune:
            # 200 trial steps: shorter trials of the shot groups rank the
            # tilings by their cold-cache starts (20 steps picked 8 x 8 at
            # 788 us per step in full runs; 200 pick 32 x 8 at 767).
            tuned_us = pl.autotune(200)
        shape = pl.shape()
        vgroups = pl.groups()
        pl.run()
        ms = pl.run()
        rec = pl.fetch()
    assert np.isfinite(rec).all()
    nsteps = 140 + nt
    bstep = vti_bytes_per_step(nx, nz, kw["pml"])
    s = ms / 1e3
    out = {"workload": f"VTI qP-acoustic (p, q; 2nd order in time, 10th order in space, split PML), "
                       f"vti_qa_2d_modelling_new.par survey: {nshot} shots of {nx}x{nz} (600 window columns "
                       f"+ 50-cell PML left, right, below), dt=0.5 ms, nt={nt} (+140 delay steps), 30 Hz",
           "shots": nshot, "shot_groups": vgroups, "cells_per_shot": nx * nz, "steps": nsteps,
           "ms_per_survey": round(ms, 1),
           "shots_per_s": round(nshot / s, 2),
           "Mcell_per_s": round(nx * nz * nshot * nsteps / s / 1e6, 1),
           "us_per_step": round(ms * 1e3 / nsteps, 2),
           "bytes_per_shot_step": bstep,
           "achieved_GBs": round(bstep * nshot * nsteps / s / 1e9, 1),
           "frac_of_8TBs": round(bstep * nshot * nsteps / s / 8e12, 4),
           "launch_shape": {"kernel": "one fused launch per step for the shots of a group", "tile_cols": shape[0],
                            "waves": shape[1], "autotuned_us_per_step": None if tuned_us is None else round(tuned_us, 2)}}
    out["traffic"] = load_traffic("vti_step_card_traffic.json", "hbm_bytes_per_shot_step",
                                  f"vti_step_kernel<{shape[0]}, {shape[1]}>")
    if not no_cpu:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        V1, E1, D1, kw1 = W.vti_card_survey(nshot=1, nt=nt_err, first=150)
        with sfd.VtiPlan(V1, E1, D1, **kw1) as pl:
            pl.run()
            got = pl.fetch()[0]
        w = oracle.vti_wavelet(kw1["dt"], kw1["f0"])
        args = dict(src=kw1["src"], rec=kw1["rec"], nsteps=140 + nt_err, delay=140, dx=kw1["dx"], dz=kw1["dz"],
                    dt=kw1["dt"], pml=kw1["pml"], wavelet=w, nthreads=cores)
        o64 = oracle.vti(V1[0], E1[0], D1[0], precision="f64", **args)
        o32 = oracle.vti(V1[0], E1[0], D1[0], precision="f32", **args)
        if kw1.get("zero_first_sample"):  # the driver's recording rule, applied by VtiPlan.fetch
            o64[:, 0] = 0.0
            o32[:, 0] = 0.0
        nrm = np.linalg.norm(o64)
        out["fp64_rel_l2"] = float(np.linalg.norm(got - o64) / nrm)
        out["fp32_oracle_vs_fp64_rel_l2"] = float(np.linalg.norm(o32 - o64) / nrm)
        out["error_sample"] = f"first {nt_err} recorded samples ({args['nsteps']} steps) of shot 150"
        # CPU baseline: one whole shot of the survey through the FP32 oracle
        full = dict(args, nsteps=nsteps, delay=140)
        t0 = time.perf_counter()
        oracle.vti(V1[0], E1[0], D1[0], precision="f32", **full)
        dt32 = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(nx * nz * nsteps / dt32 / 1e6, 2), "unit": "Mcell-updates/s",
                               "cores": cores, "kind": "port", "shots_per_s": round(1 / dt32, 4),
                               "sample": f"one whole shot ({nsteps} steps) of the survey, FP32 oracle (the "
                                         f"reference's split PML parts), OpenMP over x columns, {dt32:.1f} s"}
    return out


def srm_leg(nshot, ntrace, lt, no_cpu):
    """Convolution SRM (de-multiple/convolution_srm/convolution_srm.par:
    401 shots x 401 traces x 4000 samples): the whole prediction with the
    gathers resident in HBM, output traces per second; the FP64 oracle (the
    reference's own algorithm: nshot receiver-trace FFTs per output trace) on
    sampled output traces as the CPU baseline and the error check."""
    import seismic_code_github_amd as sfd
    rng = np.random.default_rng(31)
    csg = rng.standard_normal((nshot, ntrace, lt), dtype=np.float32)
    crg = rng.standard_normal((sfd.srm_crg_traces(nshot, ntrace), lt), dtype=np.float32)
    with sfd.SrmPlan(csg, crg) as pl:
        pl.run()
        ms = min(pl.run() for _ in range(2))
        out = pl.fetch()
    assert np.isfinite(out).all()
    ntr = nshot * ntrace
    res = {"workload": f"convolution SRM, convolution_srm.par: {nshot} shots x {ntrace} traces x {lt} samples "
                       f"(FFT length {2 * lt}), synthetic N(0,1) gathers resident in HBM",
           "ms_per_prediction": round(ms, 2), "output_traces_per_s": round(ntr / (ms / 1e3), 1)}
    if not no_cpu:
        from oracle import srm
        check = [(0, 0), (nshot // 2, ntrace // 3), (nshot - 1, ntrace - 1)]
        want = srm.predict(csg, crg, pairs=check)
        errs = [float(np.linalg.norm(out[s, t] - want[s, t]) / np.linalg.norm(want[s, t])) for s, t in check]
        res["fp64_rel_l2_max"] = max(errs)
        res["error_sample"] = f"output traces {check}"
        # CPU baseline: the first 200 output traces of shot 1 (the reference's order)
        pairs = [(1, t) for t in range(min(200, ntrace))]
        t0 = time.perf_counter()
        srm.predict(csg, crg, pairs=pairs)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(len(pairs) / dt, 3), "unit": "output traces/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(pairs)} output traces of one shot through the FP64 oracle (numpy "
                                         f"FFTs of the {nshot} receiver traces per output trace, as the reference), "
                                         f"{dt:.1f} s"}
    return res


def gradient_leg(nt, no_cpu=False):
    """Adjoint gradient: 16 shots on 801 x 801 + 40-cell sponge, 10 Hz,
    checkpoints every 50 steps: forward + checkpointed recomputation + adjoint
    per shot (sfd_acoustic_misfit_gradient, set-up and transfers included)
    against the plain forward of the same 16 shots (one launch per shot and
    step) and the batched forward.  Timed on the first `nt` of 2500 steps
    (default: all of them); shots/s scaled to 2500."""
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    true, bg, shots = W.fwi_801(nt=nt)
    obs_shots = [sfd.AcousticShot(**{**s.__dict__, "vp": true, "_keep": []}) for s in shots]
    with sfd.AcousticBatch(obs_shots) as b:
        b.run(sfd.BATCH_GRAPH)
        d_obs = [b.fetch(i) for i in range(len(shots))]
    with sfd.AcousticBatch(shots) as b:
        b.run(sfd.BATCH_PER_SHOT)
        plain_ms = b.run(sfd.BATCH_PER_SHOT)
        graph_ms = min(b.run(sfd.BATCH_GRAPH) for _ in range(2))
    # The observed data resident in HBM, as FWI iterations keep it
    # (sfd_acoustic_misfit_gradient_dev); the same call with host data after.
    d_dev = [sfd.DeviceArray(d) for d in d_obs]
    sfd.misfit_gradient(shots[:1], d_dev[:1], vp=bg, ckpt_every=50)
    t0 = time.perf_counter()
    J, g = sfd.misfit_gradient(shots, d_dev, vp=bg, ckpt_every=50)
    grad_s = time.perf_counter() - t0
    assert np.isfinite(g).all() and J >= 0
    t0 = time.perf_counter()
    Jh, gh = sfd.misfit_gradient(shots, d_obs, vp=bg, ckpt_every=50)
    host_s = time.perf_counter() - t0
    assert Jh == J and np.array_equal(gh, g)
    # The same gradient with boundary saving (reverse steps inside the
    # sponge-free box instead of the segment recomputation).
    t0 = time.perf_counter()
    Jb, gb = sfd.misfit_gradient(shots, d_dev, vp=bg, ckpt_every=50, method="boundary")
    bs_s = time.perf_counter() - t0
    # Every divergence kept on the device (a checkpoint interval >= nt: nt
    # fields per shot, 124 GB for 16 shots over 2500 steps): the forward
    # stores them and nothing is recomputed; the same gradient bit for bit.
    t0 = time.perf_counter()
    Jk, gk = sfd.misfit_gradient(shots, d_dev, vp=bg, ckpt_every=nt)
    keep_s = time.perf_counter() - t0
    assert Jk == J and np.array_equal(gk, g)
    keep_gb = len(shots) * nt * shots[0].vp.size * 4 / 1e9  # the divergence fields (pad rows aside)
    b = max(shots[0].width, shots[0].radius)
    box = (slice(b, g.shape[0] - b), slice(b, g.shape[1] - b))
    gi, gbi = g[box].astype(np.float64), gb[box].astype(np.float64)
    bs_err = float(np.linalg.norm(gbi - gi) / max(np.linalg.norm(gi), 1e-300))
    fp64 = {}
    if not no_cpu:
        fp64 = gradient_fp64_checks(shots, d_obs, bg)
    scale = 2500.0 / nt
    n = len(shots)
    # Algorithmic bytes per shot-cell and step of the fused step with uniform
    # density (bf passed as one number): forward 28 (+ 4 for the divergence
    # in the last checkpoint segment, which is not recomputed), recomputation
    # 32 (+ the divergence) in the other segments, adjoint 40 (+ divergence
    # read, gradient read and write).
    cells = int(np.prod(shots[0].vp.shape))
    nseg = -(-nt // 50)
    bpc = 72.0 + 28.0 * (nseg - 1) / nseg
    gbs = n * cells * bpc * nt / grad_s / 1e9
    return {"workload": "16 shots, 801x801 + 40-cell sponge (881x881), R=8, 10 Hz, checkpoint every 50 steps",
            "timed_steps": nt, "misfit": J,
            "fwd_adj_shots_per_s": round(n / (grad_s * scale), 3),
            "plain_fwd_shots_per_s": round(n / (plain_ms / 1e3 * scale), 3),
            "batched_fwd_shots_per_s": round(n / (graph_ms / 1e3 * scale), 3),
            "fwd_adj_vs_plain_fwd": round((plain_ms / 1e3) / grad_s, 3),
            "bytes_per_shot_cell_step": round(bpc, 2), "achieved_GBs": round(gbs, 1),
            "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4),
            "timing": "whole sfd_acoustic_misfit_gradient_dev call, observed data resident in HBM: set-up, "
                      "residuals and the gradient's transfer to the host included",
            "host_data_shots_per_s": round(n / (host_s * scale), 3),
            "divergences_kept": {"fwd_adj_shots_per_s": round(n / (keep_s * scale), 3),
                                 "bytes_per_shot_cell_step": 72,
                                 "frac_of_8TBs": round(n * cells * 72.0 * nt / keep_s / 1e9 / HBM_PEAK_GBS, 4),
                                 "divergence_GB": round(keep_gb, 1),
                                 "note": "checkpoint interval >= nt: the forward (28 B + 4 B divergence) "
                                         "stores every step's divergence, the adjoint (40 B) reads it; "
                                         "no recomputation; bitwise the checkpointed gradient"},
            "boundary_saving": {"fwd_adj_shots_per_s": round(n / (bs_s * scale), 3),
                                "box_rel_l2_vs_recompute": bs_err,
                                "box": f"[{b}, {g.shape[0] - b}) x [{b}, {g.shape[1] - b})"},
            **fp64}


def gradient_fp64_checks(shots, d_obs, bg):
    """FP64 errors of the gradient path: (a) the misfit of one shot of the
    full-size workload (all its steps) against the FP64 forward with the
    same observed data; (b) at 1/8 size (nt = 800, 4 of the 16 shots), the
    summed gradient against the FP64 adjoint."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    cores = min(16, len(os.sched_getaffinity(0)))
    s = len(shots) // 2
    sh = shots[s]
    J1, _ = sfd.misfit_gradient([sh], [d_obs[s]], vp=bg, ckpt_every=50)
    rec, _ = sfd.acoustic_forward(sh)
    den = np.full_like(bg, 2400.0)
    kw = dict(dx=sh.dx, dz=sh.dz, dt=sh.dt, radius=sh.radius, sponge=sh.width)
    r64 = oracle.acoustic(bg, den, sh.wavelet, sh.src, sh.rec, precision="f64", nthreads=cores, **kw)["record"]
    J64 = 0.5 * float(np.sum((r64 - d_obs[s].astype(np.float64)) ** 2))
    rnorm = float(np.linalg.norm(r64 - d_obs[s].astype(np.float64)) / np.linalg.norm(r64))
    true8, bg8, shots8 = W.fwi_801(nshot=16, nt=800, scale=8)
    sel = shots8[::4]
    d8 = np.full_like(bg8, 2400.0)
    obs8 = [oracle.acoustic(true8, d8, t.wavelet, t.src, t.rec, precision="f64", nthreads=cores, dx=t.dx, dz=t.dz,
                            dt=t.dt, radius=t.radius, sponge=t.width)["record"] for t in sel]
    _, g8 = sfd.misfit_gradient(sel, [o.astype(np.float32) for o in obs8], vp=bg8, ckpt_every=50)

    def adj(i):
        t = sel[i]
        kw8 = dict(dx=t.dx, dz=t.dz, dt=t.dt, radius=t.radius, sponge=t.width)
        r = oracle.acoustic(bg8, d8, t.wavelet, t.src, t.rec, precision="f64", nthreads=1, **kw8)["record"]
        return oracle.gradient(bg8, d8, t.wavelet, t.src, t.rec, r - obs8[i], **kw8)[0]

    with ThreadPoolExecutor(max_workers=min(cores, len(sel))) as ex:
        g64 = np.sum(list(ex.map(adj, range(len(sel)))), axis=0)
    return {"fp64_rel_l2": float(np.linalg.norm(rec - r64) / np.linalg.norm(r64)),
            "misfit_fp64_rel_err": abs(J1 - J64) / J64,
            "residual_over_record": rnorm,
            "gradient_fp64_rel_l2": float(np.linalg.norm(g8 - g64) / np.linalg.norm(g64)),
            "error_sample": f"gather and misfit of shot {s} (all steps) against the FP64 forward (the misfit's "
                            f"relative error is the gather's divided by residual_over_record); gradient of 4 shots "
                            f"of the workload at 1/8 size ({bg8.shape[0]}x{bg8.shape[1]}, 800 steps) against the "
                            f"FP64 adjoint"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=800)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-launches", action="store_true",
                    help="batched leg: plain launches instead of the graph replay (kernel-trace runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched leg: skip the shot-after-shot comparison (counter runs)")
    ap.add_argument("--no-visco", action="store_true")
    ap.add_argument("--visco-err-nt", type=int, default=40)
    ap.add_argument("--visco-nt", type=int, default=4000)
    ap.add_argument("--visco-batch", type=int, default=256, help="shots of the batched visco card survey (0: none)")
    ap.add_argument("--no-elastic", action="store_true")
    ap.add_argument("--elastic-err-nt", type=int, default=100)
    ap.add_argument("--elastic-nt", type=int, default=3001)
    ap.add_argument("--elastic-batch", type=int, default=32, help="shots of the batched elastic survey (0: none)")
    ap.add_argument("--no-vti", action="store_true")
    ap.add_argument("--vti-shots", type=int, default=300)
    ap.add_argument("--vti-nt", type=int, default=10000)
    ap.add_argument("--vti-err-nt", type=int, default=600)
    ap.add_argument("--vti-tiling", default=None, help="tile columns,waves (skips the autotuner)")
    ap.add_argument("--no-srm", action="store_true")
    ap.add_argument("--srm-size", default="401,401,4000", help="nshot,ntrace,lt")
    ap.add_argument("--only", choices=["visco", "grad", "batch", "elastic", "vti", "srm", "c0"], default=None,
                    help="run one secondary leg alone (profiling) and print its JSON")
    ap.add_argument("--no-c0", action="store_true")
    ap.add_argument("--tuning-file", default=None,
                    help="load the autotuner cache from this JSON file when it exists (no tuning runs: "
                         "a profile then shows the timed shots only), else tune and save it there")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming-bandwidth probe")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (the library's default launch shapes)")
    ap.add_argument("--grad-steps", type=int, default=2500)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: shots are independent and this build measures one "
              "GPU's throughput (see DESIGN.md for how shots shard across GPUs).", file=sys.stderr)
        sys.exit(2)

    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W

    if args.copy_mib:
        import torch
        n = args.copy_mib * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mib} MiB device to device")
        del src, dst

    # The tuning cache lives in the default context, which every leg uses.
    tuned_cache = False
    if args.tuning_file and os.path.exists(args.tuning_file):
        sfd.default_context().load_tuning(args.tuning_file)
        tuned_cache = sfd.default_context().tuning_size() > 0
        log(f"tuning cache: {sfd.default_context().tuning_size()} classes from {args.tuning_file}")
    vti_tiling = tuple(int(v) for v in args.vti_tiling.split(",")) if args.vti_tiling else None
    srm_size = tuple(int(v) for v in args.srm_size.split(","))
    if args.only:
        leg = {"visco": lambda: visco_leg(args.visco_err_nt, args.visco_nt, args.no_cpu, args.visco_batch),
               "grad": lambda: gradient_leg(args.grad_steps, args.no_cpu),
               "batch": lambda: batched_leg(args.batch_steps, args.no_tune or tuned_cache, args.no_cpu,
                                            not args.no_plain, args.batch_launches),
               "c0": lambda: c0_leg(args.no_cpu, args.no_tune),
               "elastic": lambda: elastic_leg(args.elastic_err_nt, args.elastic_nt, args.no_cpu, args.elastic_batch),
               "vti": lambda: vti_leg(args.vti_shots, args.vti_nt, args.vti_err_nt, args.no_cpu,
                                      args.no_tune, vti_tiling),
               "srm": lambda: srm_leg(*srm_size, args.no_cpu)}[args.only]
        res = leg()
        if args.tuning_file and not tuned_cache:
            sfd.default_context().save_tuning(args.tuning_file)
        print(json.dumps(res), flush=True)
        return

    shot = W.smooth_random_2001x4001(nt=args.nt)
    nx, nz = shot.vp.shape
    nt = len(shot.wavelet)
    log(f"workload: {nx}x{nz} cells, nt={nt}, R={shot.radius}, {shot.boundary}")
    ceiling = None
    if not args.no_ceiling:
        ceiling = {k: round(sfd.bandwidth_probe(2 << 30, kind), 1)
                   for k, kind in (("copy_GBs", sfd.BW_COPY), ("read_GBs", sfd.BW_READ), ("write_GBs", sfd.BW_WRITE))}
        ceiling["probe"] = "hand-written streaming kernels over 2 GiB (copy counts both buffers), best of 5"
        log(f"streaming ceiling: {ceiling}")
    ctx = sfd.default_context()
    plan = sfd.AcousticPlan(shot, ctx)
    tuned_us = None
    if args.waves or args.chunk:
        plan.set_tiling(args.waves, args.chunk)
    elif not args.no_tune and not tuned_cache:
        tuned_us = plan.autotune(60)
    kind, ta, tb = plan.shape()
    log(f"launch shape: {shape_dict(kind, ta, tb, tuned_us)}")
    for i in range(args.warmup):
        ms = plan.run()
        log(f"warmup shot {i}: {ms:.1f} ms")
    # Single process: no barrier needed; plan.run() synchronizes the device.
    dev_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        dev_ms.append(plan.run())
        log(f"shot {i}: {dev_ms[-1]:.1f} ms (device)")
    wall = time.perf_counter() - t0
    rec, _ = plan.fetch()
    assert np.isfinite(rec).all(), "non-finite gather"

    cells = nx * nz * nt
    value = cells * args.steps / wall / 1e6
    avg_launch_s = (sum(dev_ms) / len(dev_ms)) / 1e3 / nt
    bytes_step = algorithmic_bytes_per_step(shot)
    achieved = bytes_step / avg_launch_s / 1e9
    traffic = load_traffic("acoustic_step_cpml_2101x4101_traffic.json", "hbm_bytes_per_launch",
                           kernel_stem(kind, ta, False))

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded smooth random velocity, Ricker source)",
        "config": {"workload": f"acoustic 2D velocity-pressure staggered grid, R={shot.radius}, "
                               f"C-PML {shot.width}, {nz - 2 * shot.width}x{nx - 2 * shot.width} (z x x) interior, "
                               f"{nx}x{nz} stored, nt={nt}, dt=0.8 ms, one shot per step",
                   "cells": nx * nz, "nt": nt, "radius": shot.radius, "boundary": shot.boundary,
                   "receivers": int(len(shot.rec[0]))},
        "shots_per_s": round(args.steps / wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "ceiling_frac": None if ceiling is None else round(achieved / ceiling["copy_GBs"], 4),
                     "unit_of_measure": {0: "one time step = velocity launch + pressure launch",
                                         1: "one time step = one launch of the fused march kernel",
                                         2: "one time step = one launch of the one-launch step"}[kind],
                     "bytes_per_step": bytes_step, "avg_step_us": round(avg_launch_s * 1e6, 2)},
        "hbm_ceiling": ceiling,
        "launch_shape": shape_dict(kind, ta, tb, tuned_us),
    }
    if not args.no_cpu:
        log("fp64 error of a full-grid prefix ...")
        out.update(c1_prefix_error(shot, args.cpu_steps))
    if not args.no_c0:
        log("c0 leg ...")
        out["c0"] = c0_leg(args.no_cpu, args.no_tune)
    if not args.no_batch:
        log("batched leg ...")
        out["batched"] = batched_leg(args.batch_steps, args.no_tune or tuned_cache, args.no_cpu)
    if not args.no_visco:
        log("visco leg ...")
        out["visco"] = visco_leg(args.visco_err_nt, args.visco_nt, args.no_cpu, args.visco_batch)
    if not args.no_elastic:
        log("elastic leg ...")
        out["elastic"] = elastic_leg(args.elastic_err_nt, args.elastic_nt, args.no_cpu, args.elastic_batch)
    if not args.no_vti:
        log("vti leg ...")
        out["vti"] = vti_leg(args.vti_shots, args.vti_nt, args.vti_err_nt, args.no_cpu, args.no_tune, vti_tiling)
    if not args.no_srm:
        log("srm leg ...")
        out["srm"] = srm_leg(*srm_size, args.no_cpu)
    if not args.no_grad:
        log("gradient leg ...")
        out["gradient"] = gradient_leg(args.grad_steps, args.no_cpu)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(shot, args.cpu_steps)
    if args.tuning_file and not tuned_cache:
        ctx.save_tuning(args.tuning_file)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
