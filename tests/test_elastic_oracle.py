"""Pins of the elastic free-surface oracle (oracle/elastic_body.h) — CPU only.

forward/yang_version_2016_2_2 is Fortran that cannot be built here (no
Fortran compiler; it also needs an MPI module).  The one output the reference
ships for it, musge_test (substagger's harmonic shear modulus for the
program's own 903 x 603 shot grid), is the golden vector of the set-up; the
time loop is pinned by what the mathematics fixes:

* every stencil of calux / caluz (orders 2..10, both staggers) differentiates
  polynomials exactly up to its order and no further, with each order on the
  rows next to the surface where the reference applies it;
* substagger's averages on models that vary along one axis only, and the
  replication of the parameters above the topography;
* the free-surface conditions in the output fields (tau_zz = 0 at and above
  the surface, the tau_xz image, vz above the surface from the previous
  step's tau_zz = 0 condition), the Gaussian source after one step;
* physics: the P-wave speed in the interior and the Rayleigh-wave speed along
  the free surface, x-mirror symmetry, linearity, FP32 / FP64 agreement.
"""
import math
import os

import numpy as np
import pytest

import oracle
from conftest import oracle_threads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()


def _homog(nx, nz, vp=2000.0, vs=1150.0, rho=2000.0):
    return (np.full((nx, nz), vp, np.float32), np.full((nx, nz), vs, np.float32),
            np.full((nx, nz), rho, np.float32))


# ---------------------------------------------------------------- stencils

def _poly_field(nx, nz, axis, deg, center):
    i = np.arange(nx if axis == "x" else nz, dtype=np.float64)
    line = ((i - center) / 2.0) ** deg
    return np.repeat(line[:, None], nz, axis=1) if axis == "x" else np.repeat(line[None, :], nx, axis=0)


def _deriv_exact(deg, center, pos):
    """d/di of ((i - center)/2)^deg at position pos (cells)."""
    if deg == 0:
        return 0.0
    return deg * ((pos - center) / 2.0) ** (deg - 1) / 2.0


# A stencil exact to degree 2M leaves only the FP32 rounding of its
# coefficients (< 1e-6 here); at degree 2M + 1 its truncation error is a
# constant of order 0.4 (M = 5) and larger for smaller M.
EXACT, INEXACT = 1e-5, 1e-2


@pytest.mark.parametrize("flag", [0, 1])
def test_calux_tenth_order_exactness(flag):
    """calux (pub:1026-1084): the 10th-order pair sum is exact to degree 10
    and not at degree 11, centred half a cell before (xflag 0) or after
    (xflag 1) the output cell."""
    nx, nz = 60, 8
    center = 30.3
    for deg in range(0, 12):
        u = _poly_field(nx, nz, "x", deg, center)
        d = oracle.ssg_deriv(u, "x", flag)
        for i in (28, 30, 32):
            pos = i - 0.5 if flag == 0 else i + 0.5
            want = _deriv_exact(deg, center, pos)
            err = abs(d[i, 3] - want)
            if deg <= 10:
                assert err < EXACT * max(1.0, abs(want)), (deg, i, d[i, 3], want)
            else:
                assert err > INEXACT, (deg, i)  # the first degree it cannot follow


@pytest.mark.parametrize("flag", [0, 1])
def test_caluz_order_reduction_rows(flag):
    """caluz (pub:1160-1337): order 2M on the row `d` cells below the
    surface, with M = 5, 4, 3, 2, 1, 1 for d = 5+, 4, 3, 2, 1, 0 (zflag 0) and
    M = 5, 4, 3, 2, 1 for d = 4+, 3, 2, 1, 0 (zflag 1); a stencil of half-width
    M is exact up to degree 2M and not at 2M + 1.  Rows above the surface are
    left untouched."""
    nx, nz = 4, 60
    top = 12
    itopo = np.full(nx, top, np.int32)
    if flag == 0:
        order_of = lambda d: 5 if d >= 5 else {4: 4, 3: 3, 2: 2, 1: 1, 0: 1}[d]  # noqa: E731
    else:
        order_of = lambda d: 5 if d >= 4 else {3: 4, 2: 3, 1: 2, 0: 1}[d]  # noqa: E731
    for k in range(top, top + 8):
        M = order_of(k - top)
        center = k + 0.3
        for deg in range(0, 2 * M + 2):
            u = _poly_field(nx, nz, "z", deg, center)
            d = oracle.ssg_deriv(u, "z", flag, itopo)
            assert np.all(d[:, :top] == 0)
            pos = k - 0.5 if flag == 0 else k + 0.5
            want = _deriv_exact(deg, center, pos)
            err = abs(d[1, k] - want)
            if deg <= 2 * M:
                assert err < EXACT * max(1.0, abs(want)), (flag, k, M, deg, d[1, k], want)
            else:
                assert err > INEXACT, (flag, k, M, deg, err)


# ---------------------------------------------------------------- substagger

def test_substagger_musge_golden():
    """musge of the reference program's own run (musge_test, 903 x 603, a
    four-layer model with a flat surface at the grid top) against the
    restatement on the same layered shear modulus.  The layer values are read
    from the fixture where substagger leaves them unaveraged (row 0 and the
    last column take musg, pub:468-473); the interfaces and their mixed rows
    must land on the same rows, every value within 2 ulp.  (The reference
    build's single-precision reciprocals are not correctly rounded, so inner
    cells of a layer read 1 ulp below the modulus.)"""
    z = np.load(os.path.join(GOLDEN, "yang_musge_columns.npz"), allow_pickle=False)
    col0, colL = z["col0"], z["col_last"]
    nx, nz = 903, 603
    layers = np.unique(colL[np.r_[0:100, 105:200, 205:300, 305:603]])
    assert layers.size == 4
    # Layer of each row from the last column (musg: exact half-sums at the
    # interface row, pure values elsewhere).
    mu = np.empty(nz, np.float32)
    for k in range(nz):
        v = colL[k]
        mu[k] = v if v in layers else layers[np.searchsorted(layers, v) - 1]
    # musg(k) = (mu(k+1) + mu(k))/2 is mixed on the last row of each layer.
    rho = np.repeat(mu[None, :], nx, axis=0)
    vs = np.ones((nx, nz), np.float32)
    vp = np.full((nx, nz), 3.0, np.float32)
    st = oracle.elastic_stagger(vp, vs, rho, np.zeros(nx, np.int32))
    g = st["musge"]
    assert (g[1:902] == g[0][None, :]).all()
    np.testing.assert_array_equal(g[902], colL)  # musg itself: exact
    rel = np.abs(g[0].astype(np.float64) - col0) / np.abs(col0)
    assert rel.max() < 2.5e-7, rel.max()
    mixed_ref = np.nonzero(np.abs(np.diff(col0.astype(np.float64))) > 1e3)[0]
    mixed_ours = np.nonzero(np.abs(np.diff(g[0].astype(np.float64))) > 1e3)[0]
    np.testing.assert_array_equal(mixed_ref, mixed_ours)


def test_substagger_axis_averages_and_replication():
    """Closed forms on one-axis models (pub:345-505): a model varying along z
    only leaves x-averages exact (busg2e = 1/rho, lambmusg = lambda + 2mu,
    lambsg = lambda) and averages busg1e over the z-pair; one varying along x
    only averages busg2e and lambsg over the x-pair.  Rows above the surface
    repeat the surface row."""
    nx, nz = 12, 16
    kz = np.linspace(1.0, 2.0, nz, dtype=np.float32)
    vp = np.repeat((2000.0 * kz)[None, :], nx, 0).astype(np.float32)
    vs = np.repeat((1000.0 * kz)[None, :], nx, 0).astype(np.float32)
    rho = np.repeat((1800.0 * kz)[None, :], nx, 0).astype(np.float32)
    st = oracle.elastic_stagger(vp, vs, rho, np.zeros(nx, np.int32))
    r64, p64, s64 = rho[0].astype(np.float64), vp[0].astype(np.float64), vs[0].astype(np.float64)
    mu = r64 * s64 ** 2
    lam = r64 * p64 ** 2 - 2 * mu
    np.testing.assert_allclose(st["busg2e"][3], 1 / r64, rtol=2e-7)
    np.testing.assert_allclose(st["lambmusg"][3], lam + 2 * mu, rtol=2e-6)
    np.testing.assert_allclose(st["lambsg"][3], lam, rtol=2e-6)
    b1 = 1 / np.r_[(r64[:-1] + r64[1:]) / 2, (r64[-2] + r64[-1]) / 2]
    np.testing.assert_allclose(st["busg1e"][3, 1:], (b1[1:] + b1[:-1]) / 2, rtol=2e-7)
    assert st["busg1e"][3, 0] == pytest.approx(b1[0], rel=2e-7)

    kx = np.linspace(1.0, 1.5, nx, dtype=np.float32)
    vpx = np.repeat((2000.0 * kx)[:, None], nz, 1).astype(np.float32)
    rhox = np.repeat((1800.0 * kx)[:, None], nz, 1).astype(np.float32)
    vsx = np.zeros((nx, nz), np.float32)
    stx = oracle.elastic_stagger(vpx, vsx, rhox, np.zeros(nx, np.int32))
    r = rhox[:, 0].astype(np.float64)
    lamx = r * vpx[:, 0].astype(np.float64) ** 2
    np.testing.assert_allclose(stx["busg2e"][:-1, 5], (1 / r[:-1] + 1 / r[1:]) / 2, rtol=2e-7)
    np.testing.assert_allclose(stx["lambsg"][:-1, 5], (lamx[:-1] + lamx[1:]) / 2, rtol=2e-6)
    assert np.all(stx["musge"] == 0)  # fluid: no shear anywhere

    itopo = np.array([0, 1, 2, 5, 7, 3, 3, 0, 4, 6, 2, 1], np.int32)
    rng = np.random.default_rng(3)
    vp2 = rng.uniform(1800, 2500, (nx, nz)).astype(np.float32)
    vs2 = (vp2 / 1.8).astype(np.float32)
    rho2 = rng.uniform(1900, 2300, (nx, nz)).astype(np.float32)
    base = oracle.elastic_stagger(vp2, vs2, rho2, np.zeros(nx, np.int32))
    top = oracle.elastic_stagger(vp2, vs2, rho2, itopo)
    for name in base:
        for i in range(nx):
            t = itopo[i]
            np.testing.assert_array_equal(top[name][i, t:], base[name][i, t:])
            np.testing.assert_array_equal(top[name][i, :t], np.full(t, base[name][i, t], np.float32))


# ---------------------------------------------------------------- time loop

def _run(vp, vs, rho, itopo, **kw):
    kw.setdefault("nthreads", oracle_threads())
    return oracle.elastic(vp, vs, rho, itopo, **kw)


def test_source_footprint_after_first_step():
    """After step 1 the stresses hold only the source (addsrc, pub:1429-1466,
    get_gauss_filter_coe, pub:1470-1501): w(1) g dt / dz^2 with g the Gaussian
    exp(-r^2 / 8) normalised over the (2 rgau - 1)^2 window, centred on the
    source cell, clipped by the grid; tau_xx gets it only when the source lies
    below the surface (pri:163-167), tau_zz also above the surface."""
    nx, nz, rgau = 40, 40, 6
    vp, vs, rho = _homog(nx, nz)
    dt, dz, f0 = 1e-3, 5.0, 25.0
    off = np.arange(-(rgau - 1), rgau)
    gw = np.exp(-(off[:, None] ** 2 + off[None, :] ** 2) / 8.0)
    gw /= gw.sum()
    for isz, top in ((3, 2), (2, 2)):
        itopo = np.full(nx, top, np.int32)
        r = _run(vp, vs, rho, itopo, src=(20, isz), rec=([5], [5]), nsteps=1, delay=0, dx=5.0, dz=dz, dt=dt,
                 f0=f0, sponge=(5, 5, 5), rgau=rgau, snap_steps=(0,))
        s = r["snaps"][0]
        x = math.pi * f0 * dt  # step it = 1, idecay = 0
        w1 = math.exp(-x * x) * (1 - 2 * x * x)
        want = np.zeros((nx, nz))
        for a, da in enumerate(off):
            for b, db in enumerate(off):
                i, k = 20 + da, isz + db
                if 0 <= i < nx and 0 <= k < nz:
                    want[i, k] = w1 * gw[a, b] * dt / dz ** 2
        np.testing.assert_allclose(s[3], want, rtol=2e-6, atol=1e-14)
        if isz > top:
            np.testing.assert_allclose(s[2], want, rtol=2e-6, atol=1e-14)
        else:
            assert np.all(s[2] == 0)
        assert np.all(s[0] == 0) and np.all(s[1] == 0) and np.all(s[4] == 0)


def _compact_ricker(n, dt, f0, t0, stop):
    t = (np.arange(1, n + 1) * dt - t0) * math.pi * f0
    w = (np.exp(-t * t) * (1 - 2 * t * t)).astype(np.float32)
    w[stop:] = 0
    return w


def test_free_surface_conditions_over_topography():
    """In the output fields of a run over topography (isoewm2dfsbc,
    pri:147-196): tau_zz = 0 on and above the surface once the source is off
    (calfsbcstrzz); tau_xz one row above the surface is minus the surface value
    and zero higher up (calfsbcstrxz; row -1 is never summed, pub:1409-1422);
    vx is zero above the surface; vz one row above it is the previous step's
    vz + lambda/(lambda + 2 mu) (vx(i+1) - vx(i)) at the surface (calfsbcv is
    applied to the split parts after the step's sums), zero in the last column
    and further up."""
    nx, nz = 90, 70
    rng = np.random.default_rng(11)
    vp = rng.uniform(1900, 2300, (nx, nz)).astype(np.float32)
    vs = (vp / 1.75).astype(np.float32)
    rho = rng.uniform(1900, 2300, (nx, nz)).astype(np.float32)
    x = np.arange(nx)
    itopo = np.clip(np.round(6 + 4 * np.sin(x / 9.0)), 0, None).astype(np.int32)
    itopo[:3] = itopo[3]
    itopo[-3:] = itopo[-4]
    dt, f0 = 5e-4, 30.0
    w = _compact_ricker(400, dt, f0, 0.04, 160)
    r = _run(vp, vs, rho, itopo, src=(45, itopo[45] + 4), rec=([10], [20]), nsteps=260, delay=0, dx=5.0, dz=5.0,
             dt=dt, f0=f0, sponge=(12, 12, 12), rgau=8, wavelet=w, snap_steps=(240, 241), precision="f64")
    st = oracle.elastic_stagger(vp, vs, rho, itopo, precision="f64")
    ratio = st["lambsg"] / st["lambmusg"]
    s0, s1 = r["snaps"][0], r["snaps"][1]
    vx, vz, sxx, szz, sxz = s1
    assert np.abs(vx).max() > 0 and np.abs(szz).max() > 0
    for i in range(nx):
        t = itopo[i]
        assert np.all(szz[i, : t + 1] == 0)
        assert np.all(vx[i, :t] == 0)
        if t >= 1:
            assert sxz[i, t - 1] == -sxz[i, t]
            assert np.all(sxz[i, : t - 1] == 0)
            assert np.all(vz[i, : t - 1] == 0)
            if i < nx - 1:
                prev = s0[1][i, t] + ratio[i, t] * (s0[0][i + 1, t] - s0[0][i, t])
                assert vz[i, t - 1] == pytest.approx(prev, rel=1e-12, abs=1e-12 * np.abs(vz).max())
            else:
                assert vz[i, t - 1] == 0
    # the boundary value really is the lagged one: it differs from the same
    # expression evaluated on this step's fields
    i = 45
    t = itopo[i]
    now = vz[i, t] + ratio[i, t] * (vx[i + 1, t] - vx[i, t])
    assert abs(now - vz[i, t - 1]) > 1e-3 * abs(now)


def _xcorr_lag(a, b, dt):
    """Delay of b after a, from the cross-correlation peak refined by a
    parabola through the three best samples."""
    n = len(a) + len(b) - 1
    nf = 1 << (n - 1).bit_length()
    c = np.fft.irfft(np.conj(np.fft.rfft(a, nf)) * np.fft.rfft(b, nf), nf)
    c = np.r_[c[-(len(a) - 1):], c[: len(b)]]
    k = int(np.argmax(c))
    y0, y1, y2 = c[k - 1], c[k], c[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    return (k - (len(a) - 1) + frac) * dt


def test_p_wave_speed():
    """An explosive source in a homogeneous solid: the vz records of two
    receivers straight below it are shifted by their distance / vp."""
    nx, nz = 160, 200
    vp, vs, rho = _homog(nx, nz, 2000.0, 1100.0, 2000.0)
    dt, f0 = 4e-4, 25.0
    rz = np.array([90, 130], np.int32)
    r = _run(vp, vs, rho, np.zeros(nx, np.int32), src=(80, 50), rec=([80, 80], rz), nsteps=700, delay=0, dx=5.0,
             dz=5.0, dt=dt, f0=f0, sponge=(30, 30, 30), rgau=6)
    a, b = r["rec_vz"][0].astype(np.float64), r["rec_vz"][1].astype(np.float64)
    # window the direct P arrival of each trace (S arrives 1.8x later)
    t = np.arange(a.size) * dt
    for tr, d in ((a, 40 * 5.0), (b, 80 * 5.0)):
        tr[np.abs(t - (d / 2000.0 + 0.04)) > 1.5 / f0] = 0
    lag = _xcorr_lag(a, b, dt)
    v = 40 * 5.0 / lag
    assert v == pytest.approx(2000.0, rel=0.015), v


def test_rayleigh_wave_speed_along_free_surface():
    """A Poisson solid (vp = sqrt(3) vs) with a flat free surface and a source
    just below it: the surface wave between two surface receivers travels at
    the Rayleigh speed 0.9194 vs (root of the Rayleigh equation for
    nu = 1/4), well apart from vs itself."""
    nx, nz = 330, 80
    vs0 = 1150.0
    vp, vs, rho = _homog(nx, nz, vs0 * math.sqrt(3.0), vs0, 2000.0)
    top = 4
    itopo = np.full(nx, top, np.int32)
    dt, f0, dx = 4e-4, 15.0, 4.0
    x0 = 60
    rx = np.array([x0 + 80, x0 + 160], np.int32)
    r = _run(vp, vs, rho, itopo, src=(x0, top + 2), rec=(rx, np.full(2, top, np.int32)), nsteps=1500, delay=0,
             dx=dx, dz=dx, dt=dt, f0=f0, sponge=(40, 40, 30), rgau=6)
    cr = 0.9194 * vs0
    t = np.arange(1500) * dt
    a, b = r["rec_vz"][0].astype(np.float64), r["rec_vz"][1].astype(np.float64)
    for tr, off in ((a, 80 * dx), (b, 160 * dx)):
        tr[np.abs(t - (off / cr + 1.0 / f0)) > 1.6 / f0] = 0
    lag = _xcorr_lag(a, b, dt)
    c = 80 * dx / lag
    assert c == pytest.approx(cr, rel=0.02), (c, cr, vs0)


def test_mirror_symmetry():
    """Flat surface, homogeneous solid, source on column c of a 2c+1 wide grid:
    node fields (tau_xx, tau_zz, vz) are even about c, half-cell fields
    (vx, tau_xz) odd about c - 1/2, before any wave reaches the sponges."""
    nx, nz, c = 121, 60, 60
    vp, vs, rho = _homog(nx, nz)
    r = _run(vp, vs, rho, np.full(nx, 3, np.int32), src=(c, 8), rec=([c], [8]), nsteps=90, delay=0, dx=5.0,
             dz=5.0, dt=5e-4, f0=30.0, sponge=(10, 10, 10), rgau=6, snap_steps=(89,))
    vx, vz, sxx, szz, sxz = r["snaps"][0]
    assert np.abs(vx[:12]).max() < 1e-9 * np.abs(vx).max()  # the wave has not reached the strips
    m = np.arange(1, 45)
    for f in (vz, sxx, szz):
        np.testing.assert_allclose(f[c + m], f[c - m], rtol=0, atol=2e-6 * np.abs(f).max())
    for f in (vx, sxz):
        np.testing.assert_allclose(f[c + m], -f[c + 1 - m], rtol=0, atol=2e-6 * np.abs(f).max())


def test_linearity_and_fp64_agreement():
    nx, nz = 70, 60
    rng = np.random.default_rng(5)
    vp = rng.uniform(1900, 2300, (nx, nz)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = rng.uniform(1900, 2300, (nx, nz)).astype(np.float32)
    itopo = (3 + (np.arange(nx) // 10) % 3).astype(np.int32)
    rx = np.arange(8, 62, dtype=np.int32)
    rz = itopo[rx] + 1
    kw = dict(src=(35, 9), rec=(rx, rz), nsteps=300, delay=20, dx=5.0, dz=5.0, dt=5e-4, f0=30.0,
              sponge=(10, 10, 10), rgau=5)
    w = _compact_ricker(300, 5e-4, 30.0, 0.04, 300)
    a = _run(vp, vs, rho, itopo, wavelet=w, **kw)
    b = _run(vp, vs, rho, itopo, wavelet=2 * w, **kw)
    # exact but for the subnormal range that the far field of a 10th-order
    # stencil starts in: its rounding differs between the runs and, added to
    # normal values, flips an occasional last bit
    for k in ("rec_vx", "rec_vz"):
        np.testing.assert_allclose(b[k], 2 * a[k], rtol=0, atol=1e-6 * np.abs(b[k]).max())
    d = _run(vp, vs, rho, itopo, wavelet=w, precision="f64", **kw)
    for k in ("rec_vx", "rec_vz"):
        rel = np.linalg.norm(a[k] - d[k]) / np.linalg.norm(d[k])
        assert rel < 2e-5, (k, rel)
    # the built-in wavelet is addsrc's Ricker delayed by `delay` steps
    e = _run(vp, vs, rho, itopo, **kw)
    t = (np.arange(1, 301) * 5e-4 - 20 * 5e-4) * math.pi * 30.0
    wr = (np.exp(-t * t) * (1 - 2 * t * t)).astype(np.float32)
    f = _run(vp, vs, rho, itopo, wavelet=wr, **kw)
    # (FP32 exp in the oracle, FP64 here: the samples differ in the last bit)
    np.testing.assert_allclose(e["rec_vz"], f["rec_vz"], rtol=0, atol=1e-5 * np.abs(f["rec_vz"]).max())


def test_parameter_file_workload_matches_golden_musge():
    """The 903 x 603 shot grid built from iso_2d_modelling_sin_test.par (the
    bench workload) gives the musge the reference program wrote for it: same
    layer rows, same padding rows, values within 2 ulp."""
    from seismic_code_github_amd import workloads as W
    z = np.load(os.path.join(GOLDEN, "yang_musge_columns.npz"), allow_pickle=False)
    shot = W.elastic_par_903x603(nt=1)
    g = oracle.elastic_stagger(shot.vp, shot.vs, shot.rho, shot.itopo)["musge"]
    assert g.shape == (903, 603)
    np.testing.assert_array_equal(g[902], z["col_last"])
    rel = np.abs(g[:902].astype(np.float64) - z["col0"][None, :]) / z["col0"][None, :]
    assert rel.max() < 2.5e-7, rel.max()


# ---------------------------------------------------------------- 'abc' mode

def test_modipapml_smoothing_is_serial_in_place():
    """gauss_filter_2d (pub:592-634) replaces each cell of rows 0..itopo, column
    by column, by the window sum of the field as it stands, so smoothed cells
    feed later ones.  Row 0 (itopo = 0, rgau = 2, weights w0, w1, wd =
    exp(-{0,1,2}/8) normalised over the 3 x 3 window, clipped at the grid
    edges, not renormalised) holds a unit impulse at column 2 over a row of
    sevens: by hand, each column's value is w1 times the previous column's
    smoothed value plus its own window's unsmoothed terms (an out-of-place
    filter would miss the w1 c(i-1) terms).  Rows below itopo stay untouched."""
    nx, nz = 6, 5
    u = np.zeros((nx, nz), np.float32)
    u[2, 0] = 1.0
    u[:, 1:] = 7.0
    out = oracle.elastic_smooth(u, np.zeros(nx, np.int32), 2)
    e = np.exp(-np.array([0.0, 1.0, 2.0]) / 8.0)
    Z = e[0] + 4 * e[1] + 4 * e[2]
    w0, w1 = e[0] / Z, e[1] / Z
    # row 1 below is 7: it enters every row-0 window through the (d, 1) weights
    wd = e[2] / Z  # diagonal weight
    c0 = 7.0 * (w1 + wd)                          # col 0: window cols -1..1 clipped; (1,0) still 0
    c1 = w1 * c0 + w1 * 1.0 + 7.0 * (w1 + 2 * wd)  # smoothed (0,0), impulse (2,0), row 1
    c2 = w1 * c1 + w0 * 1.0 + 7.0 * (w1 + 2 * wd)
    c3 = w1 * c2 + 7.0 * (w1 + 2 * wd)
    c4 = w1 * c3 + 7.0 * (w1 + 2 * wd)
    c5 = w1 * c4 + 7.0 * (w1 + wd)                # col 5: right edge clipped
    np.testing.assert_allclose(out[:, 0], [c0, c1, c2, c3, c4, c5], rtol=2e-6)
    np.testing.assert_array_equal(out[:, 1:], u[:, 1:])


def test_abc_mode_absorbs_at_the_top():
    """'abc' mode (isoewm2dabc, pri:232-386): no free surface, the npzn + 1 rows
    above the surface row damp the z parts (calpmlbclrtb, pub:851-892) and the
    update domain starts npzn rows above the surface.  A P wave sent up to the
    top comes back far weaker than from the free surface, rows above the
    domain stay zero, and FP32 agrees with FP64."""
    nx, nz = 120, 140
    vp, vs, rho = _homog(nx, nz, 2000.0, 1100.0, 2000.0)
    npzn = 30
    itopo = np.full(nx, npzn + 2, np.int32)
    kw = dict(src=(60, 80), rec=([60], [70]), nsteps=520, delay=0, dx=5.0, dz=5.0, dt=5e-4, f0=30.0,
              sponge=(25, 25, 25), rgau=5, snap_steps=(150,))
    free = _run(vp, vs, rho, itopo, **kw)
    abc = _run(vp, vs, rho, itopo, mode="abc", npzn=npzn, **kw)
    t = np.arange(520) * 5e-4
    # direct arrival at 50 m, top reflection at 2*(80-32)*5 - 50 = 430 m -> 0.215 s
    late = (t > 0.17) & (t < 0.32)
    r_free = np.abs(free["rec_vz"][0, late]).max()
    r_abc = np.abs(abc["rec_vz"][0, late]).max()
    assert r_abc < 0.1 * r_free, (r_abc, r_free)
    s = abc["snaps"][0]
    top0 = itopo[0] - npzn
    assert np.all(s[:, :, :top0] == 0)
    assert np.abs(s[1, :, top0:]).max() > 0
    d = _run(vp, vs, rho, itopo, mode="abc", npzn=npzn, precision="f64", **kw)
    rel = np.linalg.norm(abc["rec_vz"] - d["rec_vz"]) / np.linalg.norm(d["rec_vz"])
    assert rel < 2e-5, rel
