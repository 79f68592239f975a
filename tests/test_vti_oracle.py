"""Pins of the VTI qP-acoustic oracle (oracle/vti.c), CPU only.

The reference (de-multiple/convolution_srm/vti_qa_2d_c3_demo1) is Fortran with
MPI and cannot be built here, and it ships no fixtures for this path, so the
restatement is pinned by what the mathematics fixes:

* coefficient_2nd / coefficient_1st: the Taylor conditions of central
  differences, closed-form weights;
* eps = delta = 0 without PML: an independent FP64 textbook second-order
  acoustic scheme (one full symmetric 11-point kernel per axis); q stays 0;
* eps = delta (elliptic): q stays 0 and p equals the isotropic run on an x
  axis stretched by sqrt(1 + 2 delta);
* the first step's impulse response (p <- p, q <- p) in closed form, and the
  derived identity for s = p + q, s'' = vz^2 ((1 + 2 eps) s_xx + p_zz), which
  fails if any of the four couplings is dropped or mis-weighted;
* qP kinematics: phase velocity vz sqrt(1 + 2 eps) along x and vz along z;
* split PML: exact x-mirror symmetry, and an order-of-magnitude less
  boundary return than the same grid without PML;
* the shot windowing of get_currshot_parameters;
* FP32 against FP64.
"""
import math

import numpy as np
import pytest

import oracle as O

C2 = np.array([5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150])
C1 = np.array([5 / 6, -5 / 21, 5 / 84, -5 / 504, 1 / 1260])


def test_coefficients_closed_form_and_taylor():
    c2, c1 = O.vti_coefficients(2).astype(np.float64), O.vti_coefficients(1).astype(np.float64)
    np.testing.assert_allclose(c2, C2, rtol=2e-6)
    np.testing.assert_allclose(c1, C1, rtol=2e-6)
    m = np.arange(1, 6, dtype=np.float64)
    for j in range(1, 6):
        # sum c_m (f(+m) + f(-m) - 2 f) = 2 sum c_m m^(2j) / (2j)! f^(2j): h^2 f'' only
        want = 1.0 if j == 1 else 0.0
        assert abs(2 * np.sum(c2 * m ** (2 * j)) / math.factorial(2 * j) - want) < 2e-5
        # sum c_m (f(+m) - f(-m)) = 2 sum c_m m^(2j-1) / (2j-1)! f^(2j-1): h f' only
        assert abs(2 * np.sum(c1 * m ** (2 * j - 1)) / math.factorial(2 * j - 1) - want) < 2e-5
    # other sizes of the same solve
    for k in (1, 2, 3, 4):
        c = O.vti_coefficients(2, k).astype(np.float64)
        assert abs(2 * np.sum(c * np.arange(1, k + 1) ** 2) / 2 - 1) < 1e-5


def test_wavelet():
    dt, f0 = 5e-4, 30.0
    w = O.vti_wavelet(dt, f0)
    assert w.size == round(2 * 0.07 / dt) + 1
    it = np.arange(1, w.size + 1)
    x = math.pi * f0 * (it * dt - 0.07)
    np.testing.assert_allclose(w, np.exp(-x * x) * (1 - 2 * x * x), atol=2e-6)
    assert np.argmax(w) == 139


def _smooth_model(nx, nz, seed, lo, hi):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nx, nz))
    for ax in (0, 1):
        for _ in range(6):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
    a = (a - a.min()) / (a.max() - a.min())
    return (lo + (hi - lo) * a).astype(np.float32)


def _textbook(vz, wav, src, rec, nsteps, dx, dz, dt, c2):
    """FP64 second-order acoustic scheme with zero Dirichlet halo: one
    11-point symmetric kernel per axis built from the weights c2."""
    k = np.concatenate([c2[::-1], [-2 * c2.sum()], c2])
    nx, nz = vz.shape
    p0 = np.zeros((nx + 10, nz + 10))
    p1 = np.zeros_like(p0)
    v2 = np.zeros_like(p0)
    v2[5:-5, 5:-5] = vz.astype(np.float64) ** 2
    out = np.zeros((len(rec[0]), nsteps))
    for n in range(nsteps):
        if n < len(wav):
            p1[src[0] + 5, src[1] + 5] += wav[n]
        lap = np.zeros_like(p1)
        for m in range(11):
            lap[5:-5, 5:-5] += k[m] * (p1[m:m + nx, 5:-5] / dx ** 2 + p1[5:-5, m:m + nz] / dz ** 2)
        p0, p1 = p1, 2 * p1 - p0 + dt * dt * v2 * lap
        p1[:5] = p1[-5:] = 0
        p1[:, :5] = p1[:, -5:] = 0
        out[:, n] = p1[np.asarray(rec[0]) + 5, np.asarray(rec[1]) + 5]
    return out


def _rec_line(nx, row):
    return (np.arange(nx), np.full(nx, row))


def test_isotropic_equals_textbook_scheme_q_zero():
    nx, nz = 61, 47
    vz = _smooth_model(nx, nz, 1, 1500, 3000)
    zero = np.zeros_like(vz)
    wav = np.random.default_rng(2).standard_normal(7).astype(np.float32)
    src, rec = (30, 20), _rec_line(nx, 5)
    # the oracle's FP32 table (pinned above); the kernel form is independent
    # (dt enters the oracle as a float)
    want = _textbook(vz, wav, src, rec, 90, 5.0, 5.0, float(np.float32(5e-4)), O.vti_coefficients(2).astype(np.float64))
    got64, p, q = O.vti(vz, zero, zero, src=src, rec=rec, nsteps=90, wavelet=wav, precision="f64", state=True)
    assert np.linalg.norm(got64 - want) <= 1e-12 * np.linalg.norm(want)
    assert not q.any()
    got32 = O.vti(vz, zero, zero, src=src, rec=rec, nsteps=90, wavelet=wav)
    assert np.linalg.norm(got32 - want) <= 2e-6 * np.linalg.norm(want)


def test_elliptic_is_a_stretched_isotropic_run():
    nx, nz = 61, 47
    vz = _smooth_model(nx, nz, 3, 1800, 2600)
    d0 = 0.18
    full = np.full_like(vz, d0)
    zero = np.zeros_like(vz)
    wav = np.random.default_rng(4).standard_normal(5).astype(np.float32)
    src, rec = (30, 20), _rec_line(nx, 8)
    ell, p, q = O.vti(vz, full, full, src=src, rec=rec, nsteps=80, wavelet=wav, precision="f64", state=True)
    assert not q.any()
    iso = O.vti(vz, zero, zero, src=src, rec=rec, nsteps=80, wavelet=wav, precision="f64",
                dx=float(np.float32(5.0 / math.sqrt(1 + 2 * np.float32(d0)))))
    # (1 + 2 delta) / dx^2 against 1 / dx'^2: equal up to the rounding of dx'
    assert np.linalg.norm(ell - iso) <= 1e-6 * np.linalg.norm(iso)
    assert np.linalg.norm(ell - iso) >= 0 and np.abs(ell).max() > 0


def test_first_step_impulse_response():
    C2 = O.vti_coefficients(2).astype(np.float64)
    nx, nz, dx, dz, dt = 31, 33, 5.0, 4.0, 4e-4
    v, ep, de = 2000.0, 0.3, 0.1
    vz = np.full((nx, nz), v, np.float32)
    w0 = 0.75
    _, p, q = O.vti(vz, np.full_like(vz, ep), np.full_like(vz, de), src=(15, 16), rec=([0], [0]), nsteps=1,
                    dx=dx, dz=dz, dt=dt, wavelet=np.array([w0], np.float32), precision="f64", state=True)
    e = lambda x: float(np.float32(x))  # noqa: E731  (model values enter as float)
    dt = float(np.float32(dt))
    t = (dt * dt) * v * v * w0
    P, Q = np.zeros((nx, nz)), np.zeros((nx, nz))
    P[15, 16] = 2 * w0 - 2 * t * ((1 + 2 * e(de)) * C2.sum() / dx ** 2 + C2.sum() / dz ** 2)
    Q[15, 16] = -2 * t * 2 * (e(ep) - e(de)) * C2.sum() / dx ** 2
    for m in range(1, 6):
        for s in (-1, 1):
            P[15 + s * m, 16] = t * (1 + 2 * e(de)) * C2[m - 1] / dx ** 2
            P[15, 16 + s * m] = t * C2[m - 1] / dz ** 2
            Q[15 + s * m, 16] = t * 2 * (e(ep) - e(de)) * C2[m - 1] / dx ** 2
    np.testing.assert_allclose(p, P, rtol=2e-6, atol=1e-12)
    np.testing.assert_allclose(q, Q, rtol=2e-6, atol=1e-12)


def test_total_field_identity():
    """s = p + q obeys s_new = 2 s - s_old + dt^2 vz^2 ((1 + 2 eps) Dxx s + Dzz p)."""
    nx, nz, dx, dz, dt = 41, 37, 5.0, 5.0, float(np.float32(5e-4))
    vz = _smooth_model(nx, nz, 5, 1600, 2400)
    ep = _smooth_model(nx, nz, 6, 0.05, 0.3)
    de = _smooth_model(nx, nz, 7, -0.05, 0.15)
    wav = np.random.default_rng(8).standard_normal(6).astype(np.float32)
    kw = dict(src=(20, 18), rec=([0], [0]), wavelet=wav, precision="f64", state=True)
    st = [O.vti(vz, ep, de, nsteps=n, **kw)[1:] for n in (30, 31, 32)]
    s = [p + q for p, q in st]
    c2 = O.vti_coefficients(2).astype(np.float64)
    k = np.concatenate([c2[::-1], [-2 * c2.sum()], c2])

    def d2(f, ax, h):
        g = np.pad(f, 5)
        out = np.zeros_like(f)
        for m in range(11):
            out += k[m] * (g[m:m + nx, 5:-5] if ax == 0 else g[5:-5, m:m + nz])
        return out / h ** 2

    v2, e64 = vz.astype(np.float64) ** 2, ep.astype(np.float64)
    rhs = 2 * s[1] - s[0] + dt * dt * v2 * ((1 + 2 * e64) * d2(s[1], 0, dx) + d2(st[1][0], 1, dz))
    assert np.linalg.norm(s[2] - rhs) <= 1e-12 * np.linalg.norm(s[2])
    # the same identity with delta in place of eps fails: the q coupling matters
    bad = rhs + dt * dt * v2 * 2 * (de.astype(np.float64) - e64) * d2(s[1], 0, dx)
    assert np.linalg.norm(s[2] - bad) > 1e-3 * np.linalg.norm(s[2])


def _pick(tr, dt):
    i = int(np.argmax(tr))
    a, b, c = tr[i - 1], tr[i], tr[i + 1]
    return (i + 0.5 * (a - c) / (a - 2 * b + c)) * dt


def test_qp_phase_velocities():
    n, dx, dt, v, ep, de = 261, 5.0, 5e-4, 2000.0, 0.25, 0.08
    vz = np.full((n, n), v, np.float32)
    w = O.vti_wavelet(dt, 30.0)
    c = n // 2
    d1, d2 = 40, 100
    rx = [c + d1, c + d2, c, c]
    rz = [c, c, c + d1, c + d2]
    nsteps = w.size // 2 + int(1.3 * d2 * dx / v / dt)
    # isotropic disc around the source (as the reference's isotropic source
    # rows): no qSV artifact is started there; the picks lie outside it
    ii, jj = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    aniso = (ii * ii + jj * jj > 20 * 20).astype(np.float32)
    rec = O.vti(vz, ep * aniso, de * aniso, src=(c, c), rec=(rx, rz), nsteps=nsteps,
                wavelet=w, dx=dx, dz=dx, dt=dt, nthreads=8)
    vx_meas = (d2 - d1) * dx / (_pick(rec[1], dt) - _pick(rec[0], dt))
    vz_meas = (d2 - d1) * dx / (_pick(rec[3], dt) - _pick(rec[2], dt))
    assert vx_meas == pytest.approx(v * math.sqrt(1 + 2 * ep), rel=0.01)
    assert vz_meas == pytest.approx(v, rel=0.01)


def _pml_case(pml, n=(121, 91)):
    nx, nz = n
    half = _smooth_model(nx // 2 + 1, nz, 9, 1800, 2600)
    vz = np.concatenate([half, half[-2::-1]], 0)  # x-mirror symmetric
    ep = np.concatenate([(h := _smooth_model(nx // 2 + 1, nz, 10, 0.0, 0.2)), h[-2::-1]], 0)
    de = 0.5 * ep
    return vz, ep, de


def test_pml_mirror_symmetry_bitwise():
    vz, ep, de = _pml_case((30, 30, 20, 30))
    nx, nz = vz.shape
    w = O.vti_wavelet(5e-4, 30.0)
    rec = (np.arange(nx), np.full(nx, 3))
    r = O.vti(vz, ep, de, src=(nx // 2, 4), rec=rec, nsteps=400, wavelet=w, pml=(30, 30, 20, 30), nthreads=4)
    np.testing.assert_array_equal(r, r[::-1])
    assert np.abs(r).max() > 0


def test_pml_absorbs():
    """Receivers of a small grid with PML against the same receivers inside
    a grid big enough that nothing returns from its edges in the run time."""
    pw, nx, nz, big = 30, 121, 91, 200
    vz = np.full((nx, nz), 2000, np.float32)
    ep, de = np.full_like(vz, 0.2), np.full_like(vz, 0.1)
    w = O.vti_wavelet(5e-4, 30.0)
    nsteps = 900
    src = (nx // 2, nz // 2)
    rec = (np.arange(pw, nx - pw, 4), np.full(len(range(pw, nx - pw, 4)), pw + 5))
    kw = dict(wavelet=w, nsteps=nsteps, nthreads=8)
    r_pml = O.vti(vz, ep, de, src=src, rec=rec, pml=(pw, pw, pw, pw), **kw)
    r_none = O.vti(vz, ep, de, src=src, rec=rec, **kw)
    B = lambda a: np.pad(a, big, mode="edge")  # noqa: E731
    r_big = O.vti(B(vz), B(ep), B(de), src=(src[0] + big, src[1] + big), rec=(rec[0] + big, rec[1] + big), **kw)
    e_pml = np.linalg.norm(r_pml - r_big) / np.linalg.norm(r_big)
    e_none = np.linalg.norm(r_none - r_big) / np.linalg.norm(r_big)
    assert e_none > 0.5
    assert e_pml < 0.1 * e_none


def test_fp32_against_fp64_with_pml():
    vz, ep, de = _pml_case(None)
    nx, nz = vz.shape
    w = O.vti_wavelet(5e-4, 30.0)
    kw = dict(src=(40, 10), rec=(np.arange(nx), np.full(nx, 6)), nsteps=500, wavelet=w, pml=(25, 20, 5, 25),
              nthreads=4)
    r32 = O.vti(vz, ep, de, **kw)
    r64 = O.vti(vz, ep, de, precision="f64", **kw)
    # FP32 rounding of the 2nd-order-in-time recursion: 1.7e-5 measured here
    assert np.linalg.norm(r32 - r64) <= 5e-5 * np.linalg.norm(r64)


def test_window():
    rng = np.random.default_rng(11)
    nxv, nzv = 40, 30
    vz, ep, de = (rng.uniform(1, 2, (nxv, nzv)).astype(np.float32) for _ in range(3))
    nl, nr, nu, nd, nsz0 = 6, 5, 3, 7, 2
    for xmin, rng_ in ((-8, 20), (10, 25), (30, 20)):
        ovz, oep, ode = O.vti_window(vz, ep, de, xmin, rng_, (nl, nr, nu, nd), nsz0)
        nx, nz = rng_ + nl + nr - 2, nzv + nu + nd - 2
        assert ovz.shape == (nx, nz)
        cols = np.clip(np.arange(1, rng_ + 1) + xmin, 1, nxv) - 1
        inner = np.s_[nl:nl + rng_, nu:nu + nzv]
        np.testing.assert_array_equal(ovz[inner], vz[cols])
        # eps / delta: window only, then the top nsz0 + nu + 6 rows isotropic
        want_e = np.zeros((nx, nz), np.float32)
        want_e[inner] = ep[cols]
        want_e[:, :nsz0 + nu + 6] = 0
        np.testing.assert_array_equal(oep, want_e)
        want_d = np.zeros((nx, nz), np.float32)
        want_d[inner] = de[cols]
        want_d[:, :nsz0 + nu + 6] = 0
        np.testing.assert_array_equal(ode, want_d)
        # vz edge-replicated into the pads (rows first, then columns)
        np.testing.assert_array_equal(ovz[:, :nu], np.repeat(ovz[:, nu:nu + 1], nu, 1))
        np.testing.assert_array_equal(ovz[:, nu + nzv:], np.repeat(ovz[:, nu + nzv - 1:nu + nzv], nz - nu - nzv, 1))
        np.testing.assert_array_equal(ovz[:nl], np.repeat(ovz[nl:nl + 1], nl, 0))
        np.testing.assert_array_equal(ovz[nl + rng_:], np.repeat(ovz[nl + rng_ - 1:nl + rng_], nx - nl - rng_, 0))
