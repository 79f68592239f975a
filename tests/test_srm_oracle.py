"""Pins of the convolution SRM oracle (oracle/srm.py), CPU only.

The reference (de-multiple/convolution_srm/convolution_srm.cpp) needs FFTW,
which the image lacks, and ships no outputs for it; its input files are
absent too.  The restatement is pinned by what the mathematics fixes: with
traces zero-padded to 2 lt, the product of spectra is the linear convolution
of the traces, so every output trace equals a brute-force time-domain sum;
spikes map to shifted spikes of the product amplitude / (ntrace + nshot);
only the trace pairs (k, is + k) of the common-receiver block contribute.
"""
import numpy as np
import pytest

from oracle import srm


def data(nshot, ntrace, lt, seed):
    rng = np.random.default_rng(seed)
    csg = rng.standard_normal((nshot, ntrace, lt)).astype(np.float32)
    crg = rng.standard_normal((srm.crg_traces_needed(nshot, ntrace), lt)).astype(np.float32)
    return csg, crg


@pytest.mark.parametrize("nshot,ntrace,lt", [(5, 5, 24), (6, 3, 17), (3, 7, 20)])
def test_frequency_domain_equals_time_domain(nshot, ntrace, lt):
    csg, crg = data(nshot, ntrace, lt, nshot * 100 + ntrace)
    got = srm.predict(csg, crg)
    for s in range(nshot):
        for t in range(ntrace):
            want = srm.predict_time_domain(csg, crg, s, t)
            np.testing.assert_allclose(got[s, t], want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))


def test_spikes_and_which_traces_pair():
    nshot, ntrace, lt = 4, 4, 16
    csg = np.zeros((nshot, ntrace, lt), np.float32)
    crg = np.zeros((srm.crg_traces_needed(nshot, ntrace), lt), np.float32)
    s, t, k = 1, 2, 1                       # shot 1, output trace 2, shot trace 1
    csg[s, k, 3] = 2.0
    base = s * ntrace + t * nshot
    crg[base + s + k, 5] = 1.5              # its partner in the block: trace s + k
    crg[base + k, 7] = 9.0                  # not a partner (trace k of the block)
    out = srm.predict(csg, crg)
    want = np.zeros(lt)
    want[3 + 5] = 2.0 * 1.5 / (ntrace + nshot)
    np.testing.assert_allclose(out[s, t], want, atol=1e-12)
    assert np.abs(np.delete(out.reshape(-1, lt), s * ntrace + t, axis=0)).max() < 1e-12
