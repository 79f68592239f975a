"""Tile-shape autotuner (sfd_acoustic_plan_autotune / sfd_acoustic_batch_autotune)
and its JSON cache (sfd_context_save_tuning / sfd_context_load_tuning).

* The autotuner picks one of the built shapes (two-launch tiled, one-launch
  tiled, march), records it under the grid
  class, and later plans / batches of that class in the context start with it.
* Every shape gives bitwise the same result, so a tuned run equals the
  default run (acoustic, elastic, VTI); the visco-acoustic candidates (tiled
  step, march kernel) agree to rounding.
* The cache survives a save / load round trip into a fresh context; saving
  merges into an existing file; a malformed file is rejected as a whole and
  entries of another GPU architecture are ignored.
"""
import json

import numpy as np
import pytest

import seismic_code_github_amd as sfd
from seismic_code_github_amd import workloads as W

pytestmark = pytest.mark.gpu

TILED = {(8, 4), (8, 8), (16, 4), (16, 8), (16, 16), (32, 4), (32, 8), (32, 16)}
# one-launch steps (rows, waves): tiled (16, 32 rows) and streaming (chunks)
FUSED = {(32, 8), (16, 4), (16, 8), (32, 4), (128, 4), (192, 4), (64, 4), (128, 8)}


def small(boundary="cpml", **kw):
    return W.smooth_random_2001x4001(nt=120, scale=10, boundary=boundary, **kw)


@pytest.mark.parametrize("boundary", ["cpml", "sponge"])
def test_plan_autotune_keeps_results_and_caches(boundary):
    shot = small(boundary)
    print(f"case: plan autotune {boundary} {shot.vp.shape}", flush=True)
    with sfd.Context(0) as ctx:
        base = sfd.AcousticPlan(shot, ctx)
        base.run()
        r0, _ = base.fetch()
        pl = sfd.AcousticPlan(shot, ctx)
        us = pl.autotune(10)
        kind, a, b = pl.shape()
        print(f"tuned: kind {kind} ({a}, {b}) {us:.2f} us/step", flush=True)
        assert us > 0
        assert ((kind == 0 and (a, b) in TILED) or (kind == 1 and a in (1, 2, 4) and b > 0)
                or (kind == 2 and (a, b) in FUSED))
        assert ctx.tuning_size() == 1
        pl.run()
        r1, _ = pl.fetch()
        np.testing.assert_array_equal(r1, r0)
        # a later plan of the same grid class starts with the tuned shape
        again = sfd.AcousticPlan(shot, ctx)
        assert again.shape() == (kind, a, b)
        for p in (base, pl, again):
            p.close()


def test_reference_arithmetic_tunes_the_march_kernel():
    shot = W.layered_301(radius=8, nt=60)
    shot.arith = "reference"
    print("case: reference-arith autotune", flush=True)
    with sfd.Context(0) as ctx:
        pl = sfd.AcousticPlan(shot, ctx)
        pl.autotune(8)
        assert pl.shape()[0] == 1
        pl.close()


def test_cache_round_trip_merge_and_rejects(tmp_path):
    path = str(tmp_path / "tune.json")
    shot = small("cpml")
    other = W.layered_301(radius=4, nt=60)
    print("case: tuning cache round trip", flush=True)
    with sfd.Context(0) as ctx:
        pl = sfd.AcousticPlan(shot, ctx)
        pl.autotune(8)
        tuned = pl.shape()
        ctx.save_tuning(path)
        pl.close()
    with sfd.Context(0) as ctx:
        assert ctx.tuning_size() == 0
        assert sfd.AcousticPlan(shot, ctx).shape() == (0, 16, 4)  # untuned default
        ctx.load_tuning(path)
        assert ctx.tuning_size() == 1
        assert sfd.AcousticPlan(shot, ctx).shape() == tuned
        # a second class, saved into the same file, is merged with the first
        q = sfd.AcousticPlan(other, ctx)
        q.autotune(8)
        ctx.save_tuning(path)
        q.close()
    doc = json.load(open(path))
    assert doc["seisfd_tuning"] == 1 and len(doc["entries"]) == 2
    classes = {e["class"] for e in doc["entries"]}
    assert any(c.startswith("acoustic:R8:cpml") for c in classes)
    assert any(c.startswith("acoustic:R4:sponge") for c in classes)

    bad = tmp_path / "bad.json"
    foreign = tmp_path / "foreign.json"
    with sfd.Context(0) as ctx:
        for text in ('{"seisfd_tuning": 1, "entries": [', '{"seisfd_tuning": 2, "entries": []}',
                     '{"seisfd_tuning": 1, "entries": [{"arch": "x", "class": "c", "kind": 5, "a": 1, "b": 1}]}',
                     '[1, 2, 3]', ''):
            bad.write_text(text)
            with pytest.raises(sfd.SeisfdError):
                ctx.load_tuning(str(bad))
            assert ctx.tuning_size() == 0
        with pytest.raises(sfd.SeisfdError):
            ctx.load_tuning(str(tmp_path / "missing.json"))
        e = dict(doc["entries"][0], arch="gfx000")
        foreign.write_text(json.dumps({"seisfd_tuning": 1, "entries": [e]}))
        ctx.load_tuning(str(foreign))
        assert ctx.tuning_size() == 0
        # saving into the foreign file keeps its entry next to this GPU's
        pl = sfd.AcousticPlan(shot, ctx)
        pl.autotune(4)
        ctx.save_tuning(str(foreign))
        pl.close()
    arches = [e["arch"] for e in json.load(open(foreign))["entries"]]
    assert "gfx000" in arches and len(arches) == 2


@pytest.mark.parametrize("boundary", ["cpml", "sponge"])
def test_batch_autotune_keeps_results(boundary):
    _, shots = W.windowed_shots(nshot=4, nt=100, scale=10, width=20, boundary=boundary)
    print(f"case: batch autotune {boundary}", flush=True)
    with sfd.Context(0) as ctx:
        b = sfd.AcousticBatch(shots, ctx)
        b.run(sfd.BATCH_GRAPH)
        r0 = [b.fetch(s) for s in range(len(shots))]
        us = b.autotune(10)
        kind, x, y = b.shape()
        print(f"tuned: kind {kind} ({x}, {y}) {us:.2f} us/step", flush=True)
        assert ((kind == 0 and (x, y) in {(8, 4), (16, 4), (16, 8), (32, 4), (32, 8)}) or kind == 1
                or (kind == 2 and (x, y) in FUSED))
        for how in (sfd.BATCH_GRAPH, sfd.BATCH_LAUNCHES):
            b.run(how)
            for s in range(len(shots)):
                np.testing.assert_array_equal(b.fetch(s), r0[s])
        b2 = sfd.AcousticBatch(shots, ctx)
        assert b2.shape() == (kind, x, y)
        b.close()
        b2.close()


def test_elastic_autotune_keeps_results():
    shot = W.elastic_par_903x603(nt=120)
    print("case: elastic autotune", flush=True)
    with sfd.Context(0) as ctx:
        base = sfd.ElasticPlan(shot, ctx)
        base.run()
        r0 = base.fetch()
        pl = sfd.ElasticPlan(shot, ctx)
        us = pl.autotune(10)
        tx, nw = pl.shape()
        print(f"tuned: {tx} x {nw} {us:.2f} us/step", flush=True)
        assert (tx, nw) in TILED | {(8, 4), (8, 8), (32, 4)} and ctx.tuning_size() == 1
        pl.run()
        r1 = pl.fetch()
        for a, b in zip(r0, r1):
            if a is not None:
                np.testing.assert_array_equal(a, b)
        assert sfd.ElasticPlan(shot, ctx).shape() == (tx, nw)


def test_visco_autotune():
    shot = W.visco_1001x2001(nt=200, scale=8)
    print("case: visco autotune", flush=True)
    with sfd.Context(0) as ctx:
        base = sfd.ViscoPlan(shot, ctx)
        base.run()
        r0 = base.fetch()[0]
        pl = sfd.ViscoPlan(shot, ctx)
        us = pl.autotune(10)
        kind, a, b = pl.shape()
        print(f"tuned: kind {kind} ({a}, {b}) {us:.2f} us/step", flush=True)
        assert (kind == 0 and (a, b) in {(8, 8), (16, 8)}) or (kind == 1 and a in (2, 4) and b > 0)
        pl.run()
        r1 = pl.fetch()[0]
        e = float(np.linalg.norm(r1.astype(np.float64) - r0) / np.linalg.norm(r0))
        assert e <= 1e-5
        assert sfd.ViscoPlan(shot, ctx).shape() == (kind, a, b)


def test_vti_autotune_keeps_results_and_caches():
    V, E, D, kw = W.vti_card_survey(nshot=3, nt=200, first=100)
    print("case: vti autotune", flush=True)
    with sfd.Context(0) as ctx:
        base = sfd.VtiPlan(V, E, D, ctx=ctx, **kw)
        base.run()
        r0 = base.fetch()
        pl = sfd.VtiPlan(V, E, D, ctx=ctx, **kw)
        us = pl.autotune(10)
        shape = pl.shape()
        print(f"tuned: {shape} {us:.2f} us/step", flush=True)
        assert shape in {(8, 4), (8, 8), (16, 4), (16, 8), (32, 8), (32, 16), (64, 8), (64, 16)}
        assert us > 0 and ctx.tuning_size() == 1
        pl.run()
        np.testing.assert_array_equal(pl.fetch(), r0)
        assert sfd.VtiPlan(V, E, D, ctx=ctx, **kw).shape() == shape
