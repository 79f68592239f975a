"""CPU oracle: convolution surface-related multiple prediction of
de-multiple/convolution_srm/convolution_srm.cpp, restated with numpy.

TEST INFRASTRUCTURE ONLY (the checker of the test-suite and of the
benchmark's CPU-baseline leg).  Product code never imports it.

The reference (convolution_srm.cpp:7-276) reads, per shot is, the shot
gather us[ntrace][lt] (its input file is [nshot][ntrace][lt]); for every
output trace it it reads nshot traces of the common-receiver file starting at
trace is * ntrace + it * nshot (the seek of :141-143 and the reads of
:174-176), zero-pads every trace to ltt = 2 lt, FFTs all of them (FFTW,
single precision, forward = e^-i), and accumulates
    mf[f] = sum over ir1 < ntrace, ir2 < nshot with receiver[ir1] == shot[ir2]
            of usf[ir1][f] * urf[ir2][f]                           (:201-216)
with receiver[ir1] = is + ir1 (:107-108) and shot[ir2] = ir2 (the branch of
:148-167 always ends with the identity: its last entry is is + it, which is
>= nshot whenever that branch runs), divides mf by ntrace + nshot (:222-223),
inverse-FFTs it (FFTW backward, unnormalised) and writes the real part / ltt
of the first lt samples (:246-252) as output trace (is, it).

This restatement follows that loop for the (is, it) pairs asked for, in
double precision (numpy's FFT); tests/test_srm_oracle.py pins it against the
time-domain linear convolution the zero padding makes it equal to.
"""
from __future__ import annotations

import numpy as np


def crg_traces_needed(nshot: int, ntrace: int) -> int:
    """Traces of the common-receiver file the reference reads (last block:
    is = nshot - 1, it = ntrace - 1, nshot traces)."""
    return (nshot - 1) * ntrace + (ntrace - 1) * nshot + nshot


def predict(csg: np.ndarray, crg: np.ndarray, pairs=None) -> np.ndarray:
    """csg [nshot, ntrace, lt], crg [ncrg, lt] (float32).  Returns the
    predicted multiples [nshot, ntrace, lt] (float64) at `pairs` (a list of
    (is, it); None = all), zero elsewhere."""
    nshot, ntrace, lt = csg.shape
    ltt = 2 * lt
    if crg.shape[0] < crg_traces_needed(nshot, ntrace):
        raise ValueError("common-receiver file too short")
    out = np.zeros((nshot, ntrace, lt))
    if pairs is None:
        pairs = [(s, t) for s in range(nshot) for t in range(ntrace)]
    spec_cache = {}
    for s, t in pairs:
        if s not in spec_cache:
            spec_cache[s] = np.fft.fft(csg[s].astype(np.float64), n=ltt, axis=1)  # usf[ir1][f]
        usf = spec_cache[s]
        base = s * ntrace + t * nshot
        urf = np.fft.fft(crg[base:base + nshot].astype(np.float64), n=ltt, axis=1)  # urf[ir2][f]
        mf = np.zeros(ltt, dtype=np.complex128)
        for ir1 in range(ntrace):
            ir2 = s + ir1  # receiver[ir1] == shot[ir2]
            if ir2 < nshot:
                mf += usf[ir1] * urf[ir2]
        mf /= (ntrace + nshot)
        out[s, t] = np.fft.ifft(mf).real[:lt]  # numpy's ifft divides by ltt, as the reference does after
    return out


def predict_time_domain(csg: np.ndarray, crg: np.ndarray, s: int, t: int) -> np.ndarray:
    """Output trace (s, t) as the linear convolution the zero padding to 2 lt
    makes the frequency-domain product equal to (FP64, brute force)."""
    nshot, ntrace, lt = csg.shape
    base = s * ntrace + t * nshot
    acc = np.zeros(2 * lt - 1)
    for ir1 in range(ntrace):
        ir2 = s + ir1
        if ir2 < nshot:
            acc += np.convolve(csg[s, ir1].astype(np.float64), crg[base + ir2].astype(np.float64))
    return acc[:lt] / (ntrace + nshot)
