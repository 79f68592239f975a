"""Batched shots (sfd_acoustic_batch_*) on MI355X.

* Every run mode (hipGraph replay, one launch per step, shot after shot)
  reproduces, bit for bit, the gather of each shot run on its own plan: the
  batch kernel executes the same per-cell arithmetic as the single-shot one.
* A captured graph replayed several times, with other device work and host
  synchronisations between the replays, gives the same gathers each time.
* The gathers match the CPU oracle (relative L2 <= 1e-5; bitwise in the
  reference arithmetic).
"""
import numpy as np
import pytest

import oracle
import seismic_code_github_amd as sfd
from seismic_code_github_amd import workloads as W
from conftest import oracle_threads

pytestmark = pytest.mark.gpu


def small_batch(boundary="cpml", nshot=5, nt=300, arith="fast"):
    _, shots = W.windowed_shots(nshot=nshot, nt=nt, scale=10, width=20, boundary=boundary)
    for s in shots:
        s.arith = arith
    return shots


@pytest.mark.parametrize("boundary,tiling", [("cpml", "march"), ("cpml", "default"), ("sponge", "march"),
                                             ("sponge", "default"), ("cpml", "fused"), ("sponge", "fused"),
                                             ("cpml", "stream"), ("sponge", "stream"), ("cpml", "8x4"),
                                             ("sponge", "16x8"), ("cpml", "32x4"), ("sponge", "32x4"),
                                             ("cpml", "32x8")])
def test_batch_modes_equal_single_plans(ctx, boundary, tiling):
    """tiling "default": batches run the two-launch tiled step (in place;
    sponge velocities kept unabsorbed between steps; "TXxNW" names another
    built tile shape of it), "fused" the one-launch
    tiled step (ping-pong), "stream" the one-launch streaming step (chunks of
    72 rows, not a multiple of its 16-row blocks), the single plans the fused
    march kernel; still bit for bit."""
    shots = small_batch(boundary)
    print(f"case: batch modes {boundary} {tiling}, shots {[s.vp.shape for s in shots]}", flush=True)
    singles = []
    for sh in shots:
        pl = sfd.AcousticPlan(sh, ctx)
        pl.set_tiling(2, 48)
        pl.run()
        singles.append(pl.fetch()[0])
        pl.close()
    b = sfd.AcousticBatch(shots, ctx)
    if tiling == "march":
        b.set_tiling(2, 48)
    elif tiling == "fused":
        b.set_fused(16, 4)
    elif tiling == "stream":
        b.set_fused(72, 4)
    elif "x" in tiling:
        b.set_tiled(*map(int, tiling.split("x")))
    for how in (sfd.BATCH_GRAPH, sfd.BATCH_LAUNCHES, sfd.BATCH_PER_SHOT):
        b.run(how)
        for s in range(len(shots)):
            rec = b.fetch(s)
            assert np.abs(rec).max() > 0
            np.testing.assert_array_equal(rec, singles[s], err_msg=f"mode {how} shot {s}")
    b.close()


def test_graph_replays_repeat(ctx):
    shots = small_batch("cpml", nshot=4, nt=200)
    print("case: graph replays", flush=True)
    b = sfd.AcousticBatch(shots, ctx)
    b.run(sfd.BATCH_GRAPH)
    first = [b.fetch(s) for s in range(len(shots))]
    other = sfd.AcousticPlan(W.layered_301(nt=50), ctx)
    for rep in range(3):
        other.run()  # unrelated eager work on the same stream between replays
        b.run(sfd.BATCH_GRAPH)
        for s in range(len(shots)):
            np.testing.assert_array_equal(b.fetch(s), first[s], err_msg=f"replay {rep} shot {s}")
    b.close()
    other.close()


def test_batch_against_oracle(ctx):
    shots = small_batch("cpml", nshot=3, nt=400)
    print("case: batch vs oracle", flush=True)
    b = sfd.AcousticBatch(shots, ctx)
    b.run(sfd.BATCH_GRAPH)
    for s, sh in enumerate(shots):
        rec = b.fetch(s)
        den = np.full_like(sh.vp, 2400.0)
        o = oracle.acoustic(sh.vp, den, sh.wavelet, sh.src, sh.rec, dt=sh.dt, radius=sh.radius, sponge=sh.width,
                            boundary="cpml", pml_r=sh.pml_r, pml_f0=sh.pml_f0, pml_vmax=float(sh.vp.max()),
                            nthreads=oracle_threads())
        rel = np.linalg.norm(rec - o["record"]) / np.linalg.norm(o["record"])
        print(f"shot {s} {sh.vp.shape}: rel-L2 vs oracle {rel:.2e}")
        assert rel <= 1e-5
    b.close()


def test_batch_reference_arith_bitwise(ctx):
    shots = small_batch("sponge", nshot=3, nt=250, arith="reference")
    print("case: batch reference arithmetic", flush=True)
    b = sfd.AcousticBatch(shots, ctx)
    b.run(sfd.BATCH_GRAPH)
    for s, sh in enumerate(shots):
        den = np.full_like(sh.vp, 2400.0)
        o = oracle.acoustic(sh.vp, den, sh.wavelet, sh.src, sh.rec, dt=sh.dt, radius=sh.radius, sponge=sh.width,
                            nthreads=oracle_threads())
        np.testing.assert_array_equal(b.fetch(s), o["record"])
    b.close()


def test_batch_rejects_mixed_shots(ctx):
    shots = small_batch("cpml", nshot=2, nt=50)
    shots[1].radius = 4
    with pytest.raises(sfd.SeisfdError) as e:
        sfd.AcousticBatch(shots, ctx)
    assert e.value.code == sfd.SFD_EINVAL


def test_shot_groups_bitwise(ctx):
    """Shot groups (sfd_acoustic_batch_set_group_bytes): a budget small
    enough for one shot per group, a middle one and one group give the same
    gathers bit for bit, for the two-launch and the one-launch steps."""
    _, shots = W.windowed_shots(nshot=5, nt=160, scale=8, width=12)
    print("case: shot groups", flush=True)
    with sfd.AcousticBatch(shots, ctx) as b:
        for kind, a, c in (("tiled", 16, 4), ("fused", 128, 4)):
            (b.set_tiled if kind == "tiled" else b.set_fused)(a, c)
            recs = {}
            for budget in (-1, 1, 3 * (1 << 20)):
                ng = b.set_group_bytes(budget)
                b.run(sfd.BATCH_GRAPH)
                recs[ng] = [b.fetch(i) for i in range(len(shots))]
            assert 1 in recs and len(shots) in recs, list(recs)
            ref = recs[1]
            assert np.abs(ref[0]).max() > 0
            for ng, rr in recs.items():
                for i in range(len(shots)):
                    np.testing.assert_array_equal(rr[i], ref[i], err_msg=f"{kind}: {ng} groups, shot {i}")
