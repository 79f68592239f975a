"""Convolution multiple prediction on the GPU (sfd_srm_*) against the oracle.

* Random gathers of several shapes (ntrace = nshot, ntrace < nshot,
  ntrace > nshot, odd lt): every output trace within rel-L2 1e-5 of the FP64
  oracle (FP32 FFTs: 1e-7 .. 1e-6 measured).
* A spike pair lands on the one output trace, at the summed delay, with the
  product amplitude / (ntrace + nshot).
* A larger survey (48 x 48 x 1000) against the oracle on sampled traces; a
  plan run twice gives bitwise the same output.
* Too short a common-receiver array is refused.
* ntrace = nshot: the batched-GEMM path and the direct sum agree with the
  oracle; ntrace != nshot refuses the GEMM path.
"""
import numpy as np
import pytest

import seismic_code_github_amd as sfd
from oracle import srm

pytestmark = pytest.mark.gpu


def data(nshot, ntrace, lt, seed):
    rng = np.random.default_rng(seed)
    csg = rng.standard_normal((nshot, ntrace, lt)).astype(np.float32)
    crg = rng.standard_normal((srm.crg_traces_needed(nshot, ntrace), lt)).astype(np.float32)
    return csg, crg


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nshot,ntrace,lt", [(7, 7, 50), (9, 4, 33), (4, 9, 64), (1, 1, 8)])
def test_against_oracle(nshot, ntrace, lt):
    csg, crg = data(nshot, ntrace, lt, 10 * nshot + ntrace)
    print(f"case: srm {nshot}x{ntrace}x{lt}", flush=True)
    got = sfd.srm_predict(csg, crg)
    want = srm.predict(csg, crg)
    e = rel(got, want)
    print(f"rel-L2 vs FP64 {e:.2e}", flush=True)
    assert e <= 1e-5
    for s in range(nshot):
        for t in range(ntrace):
            assert rel(got[s, t], want[s, t]) <= 1e-5


def test_spike():
    nshot, ntrace, lt = 5, 5, 40
    csg = np.zeros((nshot, ntrace, lt), np.float32)
    crg = np.zeros((srm.crg_traces_needed(nshot, ntrace), lt), np.float32)
    s, t, k = 2, 3, 1
    csg[s, k, 6] = 2.0
    crg[s * ntrace + t * nshot + s + k, 9] = -1.5
    print("case: srm spike", flush=True)
    got = sfd.srm_predict(csg, crg)
    want = np.zeros_like(got)
    want[s, t, 15] = 2.0 * -1.5 / (ntrace + nshot)
    np.testing.assert_allclose(got, want, atol=1e-6)


def test_larger_survey_and_rerun():
    nshot, ntrace, lt = 48, 48, 1000
    csg, crg = data(nshot, ntrace, lt, 5)
    pairs = [(0, 0), (0, 47), (13, 20), (30, 17), (47, 0), (47, 47), (24, 24)]
    print("case: srm 48x48x1000", flush=True)
    with sfd.SrmPlan(csg, crg) as pl:
        pl.run()
        a = pl.fetch()
        pl.run()
        b = pl.fetch()
    np.testing.assert_array_equal(a, b)
    want = srm.predict(csg, crg, pairs=pairs)
    for s, t in pairs:
        assert rel(a[s, t], want[s, t]) <= 1e-5


def test_refuses_short_crg():
    csg, crg = data(4, 4, 16, 1)
    with pytest.raises(sfd.SeisfdError, match="common-receiver"):
        sfd.srm_predict(csg, crg[:-1])


@pytest.mark.parametrize("n,lt", [(7, 50), (33, 120), (1, 8)])
def test_gemm_and_direct_paths(n, lt):
    """ntrace = nshot: the banded batched-GEMM path and the direct sum both
    agree with the oracle; ntrace != nshot refuses the GEMM path."""
    csg, crg = data(n, n, lt, 70 + n)
    want = srm.predict(csg, crg)
    print(f"case: srm methods {n}x{n}x{lt}", flush=True)
    outs = {}
    with sfd.SrmPlan(csg, crg) as pl:
        assert pl.method() == sfd.SRM_GEMM
        for m in (sfd.SRM_GEMM, sfd.SRM_DIRECT):
            pl.set_method(m)
            assert pl.method() == m
            pl.run()
            outs[m] = pl.fetch()
            print(f"method {m}: rel-L2 vs FP64 {rel(outs[m], want):.2e}", flush=True)
            assert rel(outs[m], want) <= 1e-5
    csg2, crg2 = data(5, 3, 16, 2)
    with sfd.SrmPlan(csg2, crg2) as pl:
        assert pl.method() == sfd.SRM_DIRECT
        with pytest.raises(sfd.SeisfdError, match="ntrace = nshot"):
            pl.set_method(sfd.SRM_GEMM)


@pytest.mark.parametrize("n,lt", [(70, 24), (260, 16)])
def test_blocked_gemm(n, lt):
    """Sizes where the GEMM path splits the shot rows into 2 and 4 blocks:
    sampled output traces (block edges, first and last shot and trace)
    against the oracle, and the whole output against the direct sum."""
    csg, crg = data(n, n, lt, n)
    q = n // 4
    pairs = [(0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1), (q - 1, n - 1), (q, 0), (n // 2, n // 2),
             (n // 2 - 1, 1), (3 * q, n - 2), (3 * q - 1, 5)]
    print(f"case: srm blocked gemm {n}x{n}x{lt}", flush=True)
    with sfd.SrmPlan(csg, crg) as pl:
        pl.set_method(sfd.SRM_GEMM)
        pl.run()
        g = pl.fetch()
        pl.set_method(sfd.SRM_DIRECT)
        pl.run()
        d = pl.fetch()
    want = srm.predict(csg, crg, pairs=pairs)
    for s, t in pairs:
        assert rel(g[s, t], want[s, t]) <= 1e-5, (s, t)
    assert rel(g, d) <= 1e-5
