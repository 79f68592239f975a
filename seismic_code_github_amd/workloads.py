"""Deterministic synthetic shots for the acoustic propagator.

The same arrays feed the HIP library and the CPU oracle.  Sizes, models and
geometries:

* ``reference_program`` — the reference program's own constant block,
  forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:28-45,137-216: 100x100
  interior + 20-cell sponge, Nt=2000, dt=0.5 ms, dx=dz=10 m, Vp=2000, DEN=2400,
  28 Hz Ricker (amplitude 50, delay 1/f), source at the centre, receivers on
  the vertical line x = Nx/2 through the source (:376-377), R=8 (C1..C8).
* ``layered_301`` — 301x301 interior + 40-cell sponge (381x381), dx=dz=10 m,
  dt=1 ms, nt=2000, three layers 1500/2500/3500 m/s with interfaces at interior
  z=100 and z=200, constant density 2400, 25 Hz Ricker at interior (z=2,
  x=150), 301 receivers at interior z=2; radius 4 (8th order) by default.
* ``visco_1001x2001`` — visco-acoustic (vs = 0): 1001 (z) x 2001 (x) model
  + 40-cell cosine sponge (the reference's npxn = npxp = npzn = npzp = 40, so
  1081 x 2081 stored), dx = dz = 10 m, dt = 1 ms, nt = 4000 recorded samples
  (+ the reference's 60-step delay), vp = seeded smooth noise (seed 2,
  sigma = 20 cells) mapped to 1500-4000 m/s, Q = seeded smooth noise (seed 3)
  mapped to 20-200, density 2400, 3 relaxation mechanisms, 20 Hz source
  (the reference's built-in wavelet and 1.5-cell Gaussian footprint) at
  model (z=2, x=1000), 2001 receivers at model z=2, R = 5 (a100..a104).
* ``elastic_par_903x603`` — the elastic free-surface program's own parameter
  file, forward/yang_version_2016_2_2/iso_2d_modelling_sin_test.par: a
  501 x 401 model at dx = dz = 5 m, one shot at x = 1250 m on the surface,
  offsets -1250..1250 m, 1000 m sponges left, right and below (npxn = npxp =
  npzp = 201) and npzn = 1 above, so the padded shot grid is nxea x nzea =
  903 x 603; nt = 3001 recorded samples after idecay = 48 steps (decay 0.048 s),
  dt = 1 ms, f0 = 25 Hz, Gaussian source window rgau = 50, receivers on every
  window column one row below the surface.  The model is the four-layer model
  whose harmonic shear modulus the program wrote (musge_test: mu = 5.625e8,
  8.1e8, 1.1025e9, 1.44e9 Pa below model rows 0, 101, 201, 301), taken as
  rho = 2250 kg/m^3 and vs = 500, 600, 700, 800 m/s, with vp = 2 vs; flat
  topography (surface on padded row 0).
* ``vti_card_survey`` — the VTI qP-acoustic program's own parameter card,
  de-multiple/convolution_srm/vti_qa_2d_c3_demo1/vti_qa_2d_modelling_new.par:
  a 300 x 300 model at dx = dz = 5 m, 300 shots from x = 0 every 5 m, offsets
  -1500 .. 1495 m (600 window columns), PML 250 m (50 cells) left, right and
  below and none above, so the padded shot grid is 698 x 348; nt = 10000
  recorded samples after the 140 delay steps of the 281-sample wavelet
  (dt = 0.5 ms, f0 = 30 Hz); source on the top row at the window centre
  (column 349), receivers on row 6 of every window column (the card's rz0 = 0
  is above row 6; the reference then leaves the first sample 0).  The card's
  model files are absent: vz, eps and delta / eps are seeded smooth noise
  (seeds 12, 13, 14, sigma = 15 cells) mapped to 1800-3500 m/s, 0-0.25 and
  0.2-0.6 (delta <= eps: the pseudo-acoustic system is unstable where
  eps < delta), each shot windowed by get_currshot_parameters (sfd.vti_window).
* ``smooth_random_2001x4001`` — 2001 (z) x 4001 (x) interior + 50-cell C-PML
  (4101 x 2101 stored), dx=dz=10 m, dt=0.8 ms, nt=6000, velocity = seeded
  Gaussian-smoothed noise (sigma = 20 cells) mapped to 1500-4500 m/s, 20 Hz
  Ricker at interior (z=2, x=2000), 4001 receivers at interior z=2, R=8.

Ricker wavelets follow the reference's formula (:199-209): t = k dt in float,
amp (1 - 2 pi^2 f^2 (t - 1/f)^2) exp(-pi^2 f^2 (t - 1/f)^2) with pi = 3.1415926,
double arithmetic, stored as float.
"""
from __future__ import annotations

import numpy as np

from . import AcousticShot, ElasticShot, ViscoShot, FIELD_P, FIELD_VX, FIELD_VZ, window_model, vti_window

PI_REF = 3.1415926


def ricker(nt: int, dt: float, freq: float, amp: float = 50.0) -> np.ndarray:
    """Acoustic_1order_FD.cpp:199-209 evaluated the same way (float t, float t1)."""
    dt32 = np.float32(dt)
    t = dt32 * np.arange(nt, dtype=np.float32)          # float t = dt*i
    t1 = np.float32(1.0 / np.float32(freq))              # float t1 = 1.0/freq
    d = (t - t1).astype(np.float64)                      # float subtraction, then promoted
    f = np.float64(np.float32(freq))
    a = 1 - 2 * PI_REF * PI_REF * f * f * d * d
    e = np.exp(-PI_REF * PI_REF * f * f * d * d)
    return (np.float64(amp) * a * e).astype(np.float32)


def smooth_random_model(nx: int, nz: int, sigma: float, vmin: float, vmax: float, seed: int) -> np.ndarray:
    """Seeded N(0,1) noise, separable Gaussian filter (sigma cells, truncated at
    4 sigma, reflecting edges), min-max mapped to [vmin, vmax]."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nx, nz), dtype=np.float64)
    sm = gaussian_filter(noise, sigma=sigma, mode="reflect", truncate=4.0)
    lo, hi = sm.min(), sm.max()
    return (vmin + (sm - lo) * (vmax - vmin) / (hi - lo)).astype(np.float32)


def reference_program(**overrides) -> AcousticShot:
    NX = NZ = 100
    PML = 20
    Nx, Nz = NX + 2 * PML, NZ + 2 * PML
    nt = 2000
    dt = 0.0005
    vp = np.full((Nx, Nz), 2000.0, dtype=np.float32)
    den = np.full((Nx, Nz), 2400.0, dtype=np.float32)
    sx, sz = Nx // 2, Nz // 2
    rz = np.arange(PML, Nx - PML, dtype=np.int32)
    rx = np.full_like(rz, sx)
    kw = dict(vp=vp, den=den, wavelet=ricker(nt, dt, 28.0), src=(sx, sz), rec=(rx, rz),
              dx=10.0, dz=10.0, dt=dt, radius=8, boundary="sponge", width=PML,
              snap_steps=(500,), snap_fields=FIELD_VX)
    kw.update(overrides)
    return AcousticShot(**kw)


def layered_301(radius: int = 4, nt: int = 2000, **overrides) -> AcousticShot:
    nb = 40
    n = 301 + 2 * nb
    z = np.arange(n) - nb
    col = np.where(z < 100, 1500.0, np.where(z < 200, 2500.0, 3500.0)).astype(np.float32)
    vp = np.ascontiguousarray(np.broadcast_to(col[None, :], (n, n)), dtype=np.float32)
    dt = 0.001
    rx = np.arange(nb, nb + 301, dtype=np.int32)
    rz = np.full_like(rx, nb + 2)
    kw = dict(vp=vp, den=None, wavelet=ricker(nt, dt, 25.0), src=(nb + 150, nb + 2), rec=(rx, rz),
              dx=10.0, dz=10.0, dt=dt, radius=radius, boundary="sponge", width=nb,
              snap_steps=(), snap_fields=FIELD_P | FIELD_VX | FIELD_VZ)
    kw.update(overrides)
    return AcousticShot(**kw)


def smooth_random_2001x4001(nt: int = 6000, radius: int = 8, seed: int = 1, scale: int = 1,
                            **overrides) -> AcousticShot:
    """``scale`` > 1 shrinks the interior by that factor (same physics per cell)
    for parity tests at oracle-friendly sizes."""
    nb = 50
    nzi, nxi = 2000 // scale + 1, 4000 // scale + 1
    nx, nz = nxi + 2 * nb, nzi + 2 * nb
    vp = smooth_random_model(nx, nz, 20.0 / scale, 1500.0, 4500.0, seed)
    dt = 0.0008
    rx = np.arange(nb, nb + nxi, dtype=np.int32)
    rz = np.full_like(rx, nb + 2)
    kw = dict(vp=vp, den=None, wavelet=ricker(nt, dt, 20.0), src=(nb + nxi // 2, nb + 2), rec=(rx, rz),
              dx=10.0, dz=10.0, dt=dt, radius=radius, boundary="cpml", width=nb, pml_r=1e-3,
              pml_f0=20.0, snap_steps=(), snap_fields=FIELD_P)
    kw.update(overrides)
    return AcousticShot(**kw)


def windowed_shots(nshot: int = 64, nt: int = 3000, scale: int = 1, radius: int = 8, width: int = 50,
                   model_seed: int = 4, window_seed: int = 5, boundary: str = "cpml"):
    """Many shots on aperture windows of one model: a 1001 (z) x 6001 (x)
    seeded smooth velocity model (sigma = 20 cells, 1500-4500 m/s); shot s
    uses a window of W_s ~ U{1000..2500} columns at a seeded origin where it
    fits, padded by `width` absorbing cells on every side (edge replication,
    as get_currshot_parameters does), source at the window centre 2 cells
    below the top of the model, receivers in every window column at that
    depth.  dx = dz = 10 m, dt = 1 ms, 20 Hz Ricker.  scale > 1 shrinks the
    model and the windows by that factor."""
    mnz, mnx = 1000 // scale + 1, 6000 // scale + 1
    model = smooth_random_model(mnx, mnz, 20.0 / scale, 1500.0, 4500.0, model_seed)
    rng = np.random.default_rng(window_seed)
    dt = 0.001
    w = ricker(nt, dt, 20.0)
    shots = []
    for s in range(nshot):
        ws = int(rng.integers(1000 // scale, 2500 // scale + 1))
        x0 = int(rng.integers(0, mnx - ws + 1))
        vp = window_model(model, x0, ws, (width, width), (width, width))
        rx = np.arange(width, width + ws, dtype=np.int32)
        rz = np.full_like(rx, width + 2)
        shots.append(AcousticShot(vp=vp, den=None, wavelet=w, src=(width + ws // 2, width + 2), rec=(rx, rz),
                                  dx=10.0, dz=10.0, dt=dt, radius=radius, boundary=boundary, width=width,
                                  pml_r=1e-3, pml_f0=20.0))
    return model, shots


def cell_updates(shot: AcousticShot) -> int:
    """Cells advanced per shot: every stored cell (boundary layers included)
    times the number of steps."""
    nx, nz = np.asarray(shot.vp).shape
    nt = shot.nt if shot.nt is not None else np.asarray(shot.wavelet).shape[0]
    return nx * nz * nt


def visco_1001x2001(nt: int = 4000, nmech: int = 3, scale: int = 1, vp_seed: int = 2, q_seed: int = 3,
                    **overrides) -> ViscoShot:
    """``scale`` > 1 shrinks the model by that factor (same physics per cell)
    for parity tests at oracle-friendly sizes; the sponge keeps 40 cells unless
    overridden."""
    nb = overrides.pop("nb", 40)
    nzi, nxi = 1000 // scale + 1, 2000 // scale + 1
    nx, nz = nxi + 2 * nb, nzi + 2 * nb
    vp = smooth_random_model(nx, nz, 20.0 / scale, 1500.0, 4000.0, vp_seed)
    qp = smooth_random_model(nx, nz, 20.0 / scale, 20.0, 200.0, q_seed)
    rx = np.arange(nb, nb + nxi, dtype=np.int32)
    rz = np.full_like(rx, nb + 2)
    kw = dict(vp=vp, qp=qp, rho=None, src=(nb + nxi // 2, nb + 2), rec=(rx, rz), nt=nt, dx=10.0, dz=10.0,
              dt=0.001, f0=20.0, radius=5, nmech=nmech, sponge=(nb, nb, nb, nb))
    kw.update(overrides)
    return ViscoShot(**kw)


def visco_card_survey(nshot: int = 256, nt: int = 2000, first: int = 1, elastic: bool = True,
                      offset_max: float = 50.0, nmech: int = 1):
    """Shots first .. first + nshot - 1 of the Q_forward card
    (forward/Q_forward/isoq_2d_modelling_simple.par): a 401 x 401 model,
    dx = dz = 10 m, offsets 0 .. offset_max (the card: 50 m, six window
    columns), 200 m sponges on every side (npxn = nint(200/10) + 1 = 21 cells,
    Q main:105-108), sx0 = sz0 = rz0 = 0, dsx = 10 m, nt = 2000, dt = 1 ms,
    f0 = 10 Hz, L = 1 (the reference's).  Per shot (Q main:124-129, Q
    sub:380-382, main:456-457): isx = 1 + (shot - 1), window model columns
    ixmine + w (get_currshot_parameters' i + ixmine) edge-padded by 21 cells,
    source at (npxn, npzn + 1), receivers on row npzn + 1 of every window
    column (0-based).  The card's model files are absent: a synthetic two-layer
    model (interface at row 200) with vp 2000 / 3000 m/s, vs = vp / sqrt(3)
    (0 with elastic=False), rho 2000 / 2300, qp 60 / 120, qs 40 / 80 (finite
    Q: the reference's override of Q to 1e5, main:418-419, is not applied).
    Returns the list of ViscoShot."""
    n = 401
    top = np.zeros((n, n), dtype=bool)
    top[:, :200] = True
    vp = np.where(top, 2000.0, 3000.0).astype(np.float32)
    vs = (vp / np.float32(np.sqrt(3.0))).astype(np.float32) if elastic else np.zeros_like(vp)
    rho = np.where(top, 2000.0, 2300.0).astype(np.float32)
    qp = np.where(top, 60.0, 120.0).astype(np.float32)
    qs = np.where(top, 40.0, 80.0).astype(np.float32)
    npx = npz = 21
    nxe = int(round(offset_max / 10.0)) + 1
    shots = []
    for i in range(first, first + nshot):
        ixmine = i  # isx = isx0 + (i - 1), isx0 = nint(sx0/dx) + 1 = 1; offset_min = 0
        win = [window_model(f, ixmine, nxe, (npx, npx), (npz, npz)) for f in (vp, vs, rho, qp, qs)]
        rx = np.arange(npx, npx + nxe, dtype=np.int32)
        rz = np.full_like(rx, npz + 1)
        shots.append(ViscoShot(vp=win[0], vs=win[1] if elastic else None, rho=win[2], qp=win[3],
                               qs=win[4] if elastic else None, src=(npx, npz + 1), rec=(rx, rz), nt=nt,
                               dx=10.0, dz=10.0, dt=1e-3, f0=10.0, radius=5, nmech=nmech,
                               sponge=(npx, npx, npz, npz)))
    return shots


def fwi_801(nshot: int = 16, nt: int = 2500, scale: int = 1, seed: int = 6):
    """Adjoint-gradient workload: 801 x 801 model + 40-cell sponge (881 x 881
    stored), dx = dz = 10 m, dt = 1 ms, nt = 2500, 10 Hz Ricker.  Background:
    linear gradient with depth 2000 -> 3500 m/s; true model: background + 3
    Gaussian anomalies of +-300 m/s (seeded centres, radii 3-8 % of the model
    size, signs); initial model: the background.  nshot sources evenly spaced
    at model z = 2, receivers in every model column at z = 2.  scale > 1
    shrinks the model by that factor.  Returns (true_vp, initial_vp, shots)
    with shots on the initial model."""
    nb = 40
    ni = 800 // scale + 1
    n = ni + 2 * nb
    z = np.clip(np.arange(n) - nb, 0, ni - 1) / max(ni - 1, 1)
    bg = np.ascontiguousarray(np.broadcast_to((2000.0 + 1500.0 * z)[None, :], (n, n)), dtype=np.float32)
    rng = np.random.default_rng(seed)
    xx, zz = np.meshgrid(np.arange(n) - nb, np.arange(n) - nb, indexing="ij")
    true = bg.astype(np.float64).copy()
    for _ in range(3):
        cx, cz = rng.uniform(0.2, 0.8, 2) * (ni - 1)
        rad = rng.uniform(0.03, 0.08) * ni
        sign = 1.0 if rng.random() < 0.5 else -1.0
        true += sign * 300.0 * np.exp(-((xx - cx) ** 2 + (zz - cz) ** 2) / (2 * rad * rad))
    true = true.astype(np.float32)
    dt = 0.001
    w = ricker(nt, dt, 10.0)
    rx = np.arange(nb, nb + ni, dtype=np.int32)
    rz = np.full_like(rx, nb + 2)
    xs = np.linspace(nb, nb + ni - 1, nshot).round().astype(int)
    shots = [AcousticShot(vp=bg, den=None, wavelet=w, src=(int(x), nb + 2), rec=(rx, rz), dx=10.0, dz=10.0, dt=dt,
                          radius=8, boundary="sponge", width=nb) for x in xs]
    return true, bg, shots


def elastic_par_903x603(nt: int = 3001, topo=None, shot: int = 1, **overrides) -> ElasticShot:
    """``topo`` (optional, [501] metres, >= 0) replaces the flat topography:
    itopoe = nint(topo / dz) + npzn (get_currshot_parameters, pub:269-272).
    ``shot`` is the card's 1-based shot number csno: the source moves by dsx =
    100 m per shot, isx = isx0 + nint((csno - 1) dsx / dx), and the window
    follows it, ixmine = isx + nint(ofsmin / dx) (spcalpa, pri:36-45)."""
    nxv, nzv, dx = 501, 401, 5.0
    npxn = npxp = npzp = 201
    npzn = 1
    nxe = 501
    nx, nz = nxe + npxn + npxp, nzv + npzn + npzp
    vs_l = np.empty(nzv, np.float32)
    vs_l[:101], vs_l[101:201], vs_l[201:301], vs_l[301:] = 500.0, 600.0, 700.0, 800.0
    vsm = np.repeat(vs_l[None, :], nxv, 0)
    model = {"vs": vsm, "vp": 2.0 * vsm, "rho": np.full((nxv, nzv), 2250.0, np.float32)}
    tp = np.zeros(nxv, np.float32) if topo is None else np.asarray(topo, np.float32)
    # get_currshot_parameters: window columns i = 1..nxe take model column
    # i + ixmine (clipped); rows pad by repetition.  sx0 = 1250 m, ofsmin =
    # -1250 m, dsx = 100 m.
    isx = 251 + int(np.floor((shot - 1) * 100.0 / dx + 0.5))
    ixmine = isx - 250
    cols = np.clip(np.arange(1, nxe + 1) + ixmine, 1, nxv) - 1
    grids = {}
    for k, m in model.items():
        g = np.empty((nx, nz), np.float32)
        g[npxn:npxn + nxe, npzn:npzn + nzv] = m[cols]
        g[npxn:npxn + nxe, :npzn] = g[npxn:npxn + nxe, npzn:npzn + 1]
        g[npxn:npxn + nxe, npzn + nzv:] = g[npxn:npxn + nxe, npzn + nzv - 1:npzn + nzv]
        g[:npxn] = g[npxn]
        g[npxn + nxe:] = g[npxn + nxe - 1]
        grids[k] = g
    topoe = np.empty(nx, np.float32)
    topoe[npxn:npxn + nxe] = tp[cols]
    topoe[:npxn] = topoe[npxn]
    topoe[npxn + nxe:] = topoe[npxn + nxe - 1]
    itopo = (np.floor(topoe / dx + 0.5).astype(np.int32) + npzn) - 1  # 0-based
    isxe = isx - ixmine + npxn + 1                  # spcalpa + get_currshot_parameters (1-based)
    sx = isxe - 1
    sz = 1 + (itopo[sx] + 1) - 1                   # isze = isz + itopoe(isxe), isz = 1
    rx = np.arange(npxn, npxn + nxe, dtype=np.int32)
    rz = (itopo[rx] + 1 + 1 - 1).astype(np.int32)  # itopo(i + npxn) + nrz0, nrz0 = 1
    kw = dict(vp=grids["vp"], vs=grids["vs"], rho=grids["rho"], itopo=itopo, src=(sx, sz), rec=(rx, rz), nt=nt,
              dx=dx, dz=dx, dt=1e-3, f0=25.0, sponge=(npxn, npxp, npzp), src_radius=50, delay=48)
    kw.update(overrides)
    return ElasticShot(**kw)


def elastic_par_survey(nshot: int = 16, nt: int = 3001, first: int = 1, **overrides):
    """Shots first .. first + nshot - 1 of the elastic card's survey on the
    903 x 603 shot grid (``elastic_par_903x603``), for a batch."""
    return [elastic_par_903x603(nt=nt, shot=first + i, **overrides) for i in range(nshot)]


def vti_receiver_row(nrz0: int, nu: int):
    """The VTI driver's recording rule (VTI main:469-476): the receiver row
    nrz0 + nu (1-based), or row 6 when that lies above it -- and in that case
    the first recorded sample (iit = 1) is never written, so it stays 0.
    Returns (0-based row, first sample zero)."""
    if nrz0 + nu >= 6:
        return nrz0 + nu - 1, False
    return 6 - 1, True


def vti_card_survey(nshot: int = 300, nt: int = 10000, first: int = 1, model_n: int = 300):
    """Shots first .. first + nshot - 1 of the VTI card (1-based, as the
    reference numbers them).  Returns (vz, eps, delta) [nshot, 698, 348] and
    the keyword arguments of VtiPlan (the card's receiver row lies above row
    6, so its records come from row 6 with the first sample zero)."""
    nxv = nzv = model_n
    dx = 5.0
    vz = smooth_random_model(nxv, nzv, 15, 1800.0, 3500.0, 12)
    ep = smooth_random_model(nxv, nzv, 15, 0.0, 0.25, 13)
    de = ep * smooth_random_model(nxv, nzv, 15, 0.2, 0.6, 14)  # delta <= eps: the system is unstable otherwise
    nl = nr = nd = 50                        # nint(250 / 5)
    nu = 0
    nsx0, nsz0, nrz0 = 1, 1, 1               # nint(0 / 5) + 1
    off_min, off_max, dsx = -1500.0, 1495.0, 5.0
    rng = int(round(off_max / dx)) - int(round(off_min / dx)) + 1
    shots = []
    for i in range(first, first + nshot):
        nsx = nsx0 + int(round((i - 1) * dsx / dx))
        xmin = nsx + int(round(off_min / dx))
        shots.append(vti_window(vz, ep, de, xmin, rng, (nl, nr, nu, nd), nsz0))
    V, E, D = (np.stack([s[k] for s in shots]) for k in range(3))
    sx = (nsx - xmin + nl) - 1               # currshot_nsx, the same for every shot
    sz = (nsz0 + nu) - 1
    row, zero_first = vti_receiver_row(nrz0, nu)
    rx = np.arange(nl, nl + rng, dtype=np.int32)
    kw = dict(dx=dx, dz=dx, dt=5e-4, nt=nt, f0=30.0, src=(sx, sz), rec=(rx, np.full(rng, row, np.int32)),
              pml=(nl, nr, nu, nd), zero_first_sample=zero_first)
    return V, E, D, kw
