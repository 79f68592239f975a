"""ctypes binding of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()``
and by ``bench.py``'s CPU-baseline leg, always as the checker.  The product
(``seismic_code_github_amd``) never imports it.

The oracle restates forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:126-393
in C (``acoustic.c``); see ``oracle.h`` for the argument meaning.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SEISFD_ORACLE_LIB selects another build of the same sources (the tests'
# AddressSanitizer build); the default is the in-tree liboracle.so.
LIB_PATH = os.environ.get("SEISFD_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

# Reference coefficients C1..C8 as written in Acoustic_1order_FD.cpp:38-45 (float).
REF_C8 = np.array([1.2340911, -1.0664985e-1, 2.3036367e-2, -5.3423856e-3,
                   1.0772712e-3, -1.6641888e-4, 1.7021711e-5, -8.5234642e-7],
                  dtype=np.float32)
# a80..a83 and a100..a104, yang_version_2016_2_2/Iso_multishot_fd_module_zy_m2.f90:17-26.
REF_A8 = np.array([1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0],
                  dtype=np.float32)
REF_A10 = np.array([19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0,
                    -405.0 / 229376.0, 35.0 / 294912.0], dtype=np.float32)


def coefs_for_radius(radius: int) -> np.ndarray:
    """The reference's own table for R = 8, 4 or 5; Taylor weights otherwise."""
    if radius == 8:
        return REF_C8.copy()
    if radius == 4:
        return REF_A8.copy()
    if radius == 5:
        return REF_A10.copy()
    return staggered_coefs(radius).astype(np.float32)


class _Args(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int), ("k0", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("radius", ctypes.c_int), ("coef", ctypes.POINTER(ctypes.c_float)),
        ("sponge", ctypes.c_int), ("absorb", ctypes.POINTER(ctypes.c_float)),
        ("vp", ctypes.POINTER(ctypes.c_float)), ("den", ctypes.POINTER(ctypes.c_float)),
        ("wavelet", ctypes.POINTER(ctypes.c_float)),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int),
        ("nrec", ctypes.c_int), ("rx", ctypes.POINTER(ctypes.c_int)),
        ("rz", ctypes.POINTER(ctypes.c_int)),
        ("nsnap", ctypes.c_int), ("snap_steps", ctypes.POINTER(ctypes.c_int)),
        ("nthreads", ctypes.c_int),
        ("boundary", ctypes.c_int), ("pml_r", ctypes.c_float),
        ("pml_vmax", ctypes.c_float), ("pml_f0", ctypes.c_float),
    ]


_lib = None


def build() -> None:
    """Compile liboracle.so (and the reference program when its source is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_acoustic_f32.argtypes = [ctypes.POINTER(_Args), fp, fp, fp, fp]
        L.oracle_acoustic_f64.argtypes = [ctypes.POINTER(_Args), dp, dp, dp, dp]
        L.oracle_sponge_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.oracle_ricker.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, fp]
        L.oracle_staggered_coefs.argtypes = [ctypes.c_int, dp]
        L.oracle_acoustic_gradient_f64.argtypes = [ctypes.POINTER(_Args), dp, dp, dp]
        L.oracle_window.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        L.oracle_cpml_coef.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, dp, dp]
        _lib = L
    return _lib


def _ptr(a, ctype=ctypes.c_float):
    if a is None:
        return ctypes.POINTER(ctype)()
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def sponge_table(nx: int, nz: int, width: int) -> np.ndarray:
    out = np.empty((nx, nz), dtype=np.float32)
    lib().oracle_sponge_table(nx, nz, width, _ptr(out))
    return out


def ricker(nt: int, dt: float, freq: float, amp: float = 50.0) -> np.ndarray:
    out = np.empty(nt, dtype=np.float32)
    lib().oracle_ricker(nt, dt, freq, amp, _ptr(out))
    return out


def staggered_coefs(radius: int) -> np.ndarray:
    out = np.empty(radius, dtype=np.float64)
    lib().oracle_staggered_coefs(radius, _ptr(out, ctypes.c_double))
    return out


def window(model, ixmine, nxe, npxn, npxp, npzn, npzp):
    """get_currshot_parameters' window + edge padding (oracle/window.c)."""
    m = np.ascontiguousarray(model, dtype=np.float32)
    nx, nz = m.shape
    out = np.full((nxe + npxn + npxp, nz + npzn + npzp), np.nan, dtype=np.float32)
    lib().oracle_window(_ptr(m), nx, nz, ixmine, nxe, npxn, npxp, npzn, npzp, _ptr(out))
    return out


def cpml_coef(i, shift, n, w, h, dt, R, vmax, f0):
    b = ctypes.c_double()
    a = ctypes.c_double()
    lib().oracle_cpml_coef(i, shift, n, w, h, dt, R, vmax, f0, ctypes.byref(b), ctypes.byref(a))
    return b.value, a.value


def acoustic(vp, den, wavelet, src, rec, *, dx=10.0, dz=10.0, dt=5e-4, radius=8,
             sponge=20, nt=None, k0=0, snap_steps=(), state=None, pml_state=None,
             precision="f32", nthreads=1, boundary="sponge", pml_r=1e-3,
             pml_vmax=None, pml_f0=None, coef=None, absorb=None):
    """Run the oracle.  vp, den: [nx, nz] float32 (z fastest).  src = (sx, sz);
    rec = (rx[], rz[]).  Returns dict(record [nrec, nt], snaps [nsnap, 3, nx, nz],
    state [3, nx, nz], pml [4, nx, nz] or None)."""
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    den = np.ascontiguousarray(den, dtype=np.float32)
    nx, nz = vp.shape
    wavelet = np.ascontiguousarray(wavelet, dtype=np.float32)
    if nt is None:
        nt = wavelet.shape[0] - k0
    assert wavelet.shape[0] >= k0 + nt
    coef = coefs_for_radius(radius) if coef is None else np.ascontiguousarray(coef, dtype=np.float32)
    rx = np.ascontiguousarray(rec[0], dtype=np.int32)
    rz = np.ascontiguousarray(rec[1], dtype=np.int32)
    snaps_idx = np.ascontiguousarray(sorted(snap_steps), dtype=np.int32)
    absorb_arr = None if absorb is None else np.ascontiguousarray(absorb, dtype=np.float32)
    bflag = {"sponge": 0, "cpml": 1}[boundary]
    args = _Args(nx, nz, nt, k0, dx, dz, dt, radius, _ptr(coef), sponge, _ptr(absorb_arr),
                 _ptr(vp), _ptr(den), _ptr(wavelet), int(src[0]), int(src[1]),
                 rx.size, _ptr(rx, ctypes.c_int), _ptr(rz, ctypes.c_int),
                 snaps_idx.size, _ptr(snaps_idx, ctypes.c_int), nthreads, bflag,
                 pml_r, float(pml_vmax if pml_vmax is not None else vp.max()),
                 float(pml_f0 if pml_f0 is not None else 0.0))
    dtype = np.float32 if precision == "f32" else np.float64
    ct = ctypes.c_float if precision == "f32" else ctypes.c_double
    record = np.zeros((rx.size, nt), dtype=dtype)
    snaps = np.zeros((snaps_idx.size, 3, nx, nz), dtype=dtype)
    st = np.zeros((3, nx, nz), dtype=dtype) if state is None else np.array(state, dtype=dtype, copy=True)
    pml = None
    if bflag == 1:
        pml = np.zeros((4, nx, nz), dtype=dtype) if pml_state is None else np.array(pml_state, dtype=dtype, copy=True)
    fn = lib().oracle_acoustic_f32 if precision == "f32" else lib().oracle_acoustic_f64
    rc = fn(ctypes.byref(args), _ptr(record, ct), _ptr(snaps, ct) if snaps_idx.size else ctypes.POINTER(ct)(),
            _ptr(st, ct), _ptr(pml, ct))
    if rc != 0:
        raise ValueError(f"oracle_acoustic returned {rc}")
    return {"record": record, "snaps": snaps, "state": st, "pml": pml}


def gradient(vp, den, wavelet, src, rec, seed, *, dx=10.0, dz=10.0, dt=5e-4, radius=8, sponge=20,
             coef=None, boundary="sponge", pml_r=1e-3, pml_vmax=None, pml_f0=None):
    """FP64 discrete adjoint (oracle/adjoint.c): dJ/dvp for
    J = sum(seed * record); returns (grad [nx, nz], record [nrec, nt]).
    boundary "cpml": the C-PML step of width `sponge`, its profiles fixed by
    pml_vmax (default: the model's maximum) and held constant."""
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    den = np.ascontiguousarray(den, dtype=np.float32)
    nx, nz = vp.shape
    wavelet = np.ascontiguousarray(wavelet, dtype=np.float32)
    nt = wavelet.shape[0]
    coef = coefs_for_radius(radius) if coef is None else np.ascontiguousarray(coef, dtype=np.float32)
    rx = np.ascontiguousarray(rec[0], dtype=np.int32)
    rz = np.ascontiguousarray(rec[1], dtype=np.int32)
    seed = np.ascontiguousarray(seed, dtype=np.float64)
    assert seed.shape == (rx.size, nt)
    bflag = {"sponge": 0, "cpml": 1}[boundary]
    args = _Args(nx, nz, nt, 0, dx, dz, dt, radius, _ptr(coef), sponge, _ptr(None), _ptr(vp), _ptr(den),
                 _ptr(wavelet), int(src[0]), int(src[1]), rx.size, _ptr(rx, ctypes.c_int), _ptr(rz, ctypes.c_int),
                 0, _ptr(None, ctypes.c_int), 1, bflag, pml_r if bflag else 0.0,
                 float(pml_vmax if pml_vmax is not None else vp.max()) if bflag else 0.0,
                 float(pml_f0 or 0.0) if bflag else 0.0)
    grad = np.zeros((nx, nz), dtype=np.float64)
    record = np.zeros((rx.size, nt), dtype=np.float64)
    rc = lib().oracle_acoustic_gradient_f64(ctypes.byref(args), _ptr(seed, ctypes.c_double),
                                            _ptr(grad, ctypes.c_double), _ptr(record, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle_acoustic_gradient_f64 returned {rc}")
    return grad, record


def reference_config():
    """The reference program's hard-coded configuration, Acoustic_1order_FD.cpp:28-45,137-216:
    NX=NZ=100, PML=20, Nt=2000, dx=dz=10, dt=0.5 ms, Vp=2000, DEN=2400, 28 Hz
    Ricker (amplitude 50), source at (Nx/2, Nz/2), receivers P1[Nx/2][PML..Nz-PML)."""
    NX = NZ = 100
    PML = 20
    Nx, Nz = NX + 2 * PML, NZ + 2 * PML
    nt = 2000
    dt = np.float32(0.0005)
    vp = np.full((Nx, Nz), 2000.0, dtype=np.float32)
    den = np.full((Nx, Nz), 2400.0, dtype=np.float32)
    w = ricker(nt, dt, 28.0, 50.0)
    sx, sz = Nx // 2, Nz // 2
    rz = np.arange(PML, Nx - PML, dtype=np.int32)
    rx = np.full_like(rz, sx)
    return dict(vp=vp, den=den, wavelet=w, src=(sx, sz), rec=(rx, rz), dx=10.0, dz=10.0,
                dt=float(dt), radius=8, sponge=PML, nt=nt)


class _VArgs(ctypes.Structure):
    """Mirror of ``oracle_visco_args`` (oracle.h)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nsteps", ctypes.c_int), ("it0", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float), ("f0", ctypes.c_float),
        ("radius", ctypes.c_int), ("coef", ctypes.POINTER(ctypes.c_float)),
        ("nmech", ctypes.c_int), ("q_scale", ctypes.c_float),
        ("f_low", ctypes.c_float), ("f_high", ctypes.c_float),
        ("npxn", ctypes.c_int), ("npxp", ctypes.c_int), ("npzn", ctypes.c_int), ("npzp", ctypes.c_int),
        ("sponge_r", ctypes.c_float),
        ("vp", ctypes.POINTER(ctypes.c_float)), ("vs", ctypes.POINTER(ctypes.c_float)),
        ("rho", ctypes.POINTER(ctypes.c_float)), ("qp", ctypes.POINTER(ctypes.c_float)),
        ("qs", ctypes.POINTER(ctypes.c_float)),
        ("isx", ctypes.c_int), ("isz", ctypes.c_int), ("src_sigma", ctypes.c_float),
        ("wavelet", ctypes.POINTER(ctypes.c_float)),
        ("nrec", ctypes.c_int), ("rx", ctypes.POINTER(ctypes.c_int)), ("rz", ctypes.POINTER(ctypes.c_int)),
        ("rec_delay", ctypes.c_int),
        ("nsnap", ctypes.c_int), ("snap_steps", ctypes.POINTER(ctypes.c_int)),
        ("nthreads", ctypes.c_int), ("init", ctypes.POINTER(ctypes.c_float)),
    ]


def _bind_visco(L):
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_visco_f32.argtypes = [ctypes.POINTER(_VArgs), fp, fp, fp]
    L.oracle_visco_f64.argtypes = [ctypes.POINTER(_VArgs), dp, dp, dp]
    d = ctypes.c_double
    L.oracle_visco_tau_f32.argtypes = [ctypes.c_int, d, d, d, d, d, fp, fp]
    L.oracle_visco_tau_f64.argtypes = [ctypes.c_int, d, d, d, d, d, dp, dp]
    L.oracle_rsg_deriv_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, d,
                                       ctypes.c_int, fp]
    L.oracle_csponge_f32.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, fp, fp]
    L.oracle_visco_source_weights_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_float, fp]


def _vlib():
    L = lib()
    if not hasattr(L, "_visco_bound"):
        _bind_visco(L)
        L._visco_bound = True
    return L


def rsg_deriv(u, direction, shift, h=10.0, radius=5, coef=None):
    """One rotated-grid derivative of the visco restatement (FP64):
    direction "x" (calderxabs) or "z" (calderzabs), xflag/zflag = shift."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nx, nz = u.shape
    coef = coefs_for_radius(radius) if coef is None else np.ascontiguousarray(coef, dtype=np.float32)
    out = np.zeros_like(u)
    rc = _vlib().oracle_rsg_deriv_f64({"x": 0, "z": 1}[direction], shift, nx, nz, _ptr(u, ctypes.c_double),
                                      _ptr(out, ctypes.c_double), h, radius, _ptr(coef))
    if rc != 0:
        raise ValueError(f"oracle_rsg_deriv_f64 returned {rc}")
    return out


def csponge(np_, h, r=1e-3, lo=True):
    """(node, half) tables of one csponge side, outermost cell first."""
    node = np.zeros(np_ + 1, dtype=np.float32)
    half = np.zeros(np_ + 1, dtype=np.float32)
    _vlib().oracle_csponge_f32(np_, h, r, 1 if lo else 0, _ptr(node), _ptr(half))
    return node, half


def visco_source_weights(nx, nz, isx, isz, sigma=1.5):
    """Per-cell weight of addsrc's Gaussian footprint [nx, nz]."""
    w = np.zeros((nx, nz), dtype=np.float32)
    rc = _vlib().oracle_visco_source_weights_f32(nx, nz, isx, isz, sigma, _ptr(w))
    if rc != 0:
        raise ValueError(rc)
    return w


def visco_tau(nmech, f0, q, q_scale=0.9, f_low=0.0, f_high=0.0, precision="f32"):
    """(tau_sigma [L], tau) of the tau-method (pcalpas, Q sub:506-542)."""
    L = lib()
    if not hasattr(L, "_visco_bound"):
        _bind_visco(L)
        L._visco_bound = True
    dtype, ct = (np.float32, ctypes.c_float) if precision == "f32" else (np.float64, ctypes.c_double)
    ts = np.zeros(nmech, dtype=dtype)
    tau = ct()
    fn = L.oracle_visco_tau_f32 if precision == "f32" else L.oracle_visco_tau_f64
    fn(nmech, f0, q, q_scale, f_low, f_high, _ptr(ts, ct), ctypes.byref(tau))
    return ts, tau.value


def visco(vp, rho, qp, *, src, rec, nsteps, dx=10.0, dz=10.0, dt=1e-3, f0=20.0, vs=None, qs=None,
          radius=5, coef=None, nmech=1, q_scale=0.9, f_low=0.0, f_high=0.0, sponge=(40, 40, 40, 40),
          sponge_r=1e-3, src_sigma=1.5, wavelet=None, rec_delay=0, snap_steps=(), it0=1,
          precision="f32", nthreads=1, init=None):
    """Run the visco-elastic RSG oracle (oracle/visco_body.h).  Fields
    [nx, nz] float32 (z fastest); sponge = (npxn, npxp, npzn, npzp).
    init: None (start from rest, as the reference) or [5, nx, nz] initial
    sxx, szz, sxz, vx, vz (the oracle's own plane-wave pins).
    Returns dict(rec_zz, rec_xx [nrec, nsteps - rec_delay],
    snaps [nsnap, 5, nx, nz] = sxx, szz, sxz, vx, vz)."""
    L = lib()
    if not hasattr(L, "_visco_bound"):
        _bind_visco(L)
        L._visco_bound = True
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    vp, rho, qp, vs, qs = f32(vp), f32(rho), f32(qp), f32(vs), f32(qs)
    nx, nz = vp.shape
    coef = coefs_for_radius(radius) if coef is None else f32(coef)
    rx = np.ascontiguousarray(rec[0], dtype=np.int32)
    rz = np.ascontiguousarray(rec[1], dtype=np.int32)
    steps = np.ascontiguousarray(sorted(snap_steps), dtype=np.int32)
    w = f32(wavelet)
    if w is not None:
        assert w.shape[0] >= nsteps
    ini = f32(init)
    if ini is not None:
        assert ini.shape == (5, nx, nz)
    args = _VArgs(nx, nz, nsteps, it0, dx, dz, dt, f0, radius, _ptr(coef), nmech, q_scale, f_low, f_high,
                  *sponge, sponge_r, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(qp), _ptr(qs),
                  int(src[0]), int(src[1]), src_sigma, _ptr(w), rx.size, _ptr(rx, ctypes.c_int),
                  _ptr(rz, ctypes.c_int), rec_delay, steps.size, _ptr(steps, ctypes.c_int), nthreads, _ptr(ini))
    dtype, ct = (np.float32, ctypes.c_float) if precision == "f32" else (np.float64, ctypes.c_double)
    nrt = max(nsteps - rec_delay, 0)
    rzz = np.zeros((rx.size, nrt), dtype=dtype)
    rxx = np.zeros((rx.size, nrt), dtype=dtype)
    snaps = np.zeros((steps.size, 5, nx, nz), dtype=dtype)
    fn = L.oracle_visco_f32 if precision == "f32" else L.oracle_visco_f64
    rc = fn(ctypes.byref(args), _ptr(rzz, ct), _ptr(rxx, ct), _ptr(snaps, ct) if steps.size else ctypes.POINTER(ct)())
    if rc != 0:
        raise ValueError(f"oracle_visco returned {rc}")
    return {"rec_zz": rzz, "rec_xx": rxx, "snaps": snaps}


class _EArgs(ctypes.Structure):
    """Mirror of ``oracle_elastic_args`` (oracle.h)."""
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nsteps", ctypes.c_int), ("delay", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float), ("f0", ctypes.c_float),
        ("vp", ctypes.POINTER(ctypes.c_float)), ("vs", ctypes.POINTER(ctypes.c_float)),
        ("rho", ctypes.POINTER(ctypes.c_float)), ("itopo", ctypes.POINTER(ctypes.c_int)),
        ("npxn", ctypes.c_int), ("npxp", ctypes.c_int), ("npzp", ctypes.c_int),
        ("sponge_r", ctypes.c_float), ("isx", ctypes.c_int), ("isz", ctypes.c_int), ("rgau", ctypes.c_int),
        ("wavelet", ctypes.POINTER(ctypes.c_float)),
        ("nrec", ctypes.c_int), ("rx", ctypes.POINTER(ctypes.c_int)), ("rz", ctypes.POINTER(ctypes.c_int)),
        ("nsnap", ctypes.c_int), ("snap_steps", ctypes.POINTER(ctypes.c_int)),
        ("nthreads", ctypes.c_int), ("mode", ctypes.c_int), ("npzn", ctypes.c_int),
    ]


def _elib():
    L = lib()
    if not hasattr(L, "_elastic_bound"):
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_elastic_f32.argtypes = [ctypes.POINTER(_EArgs), fp, fp, fp]
        L.oracle_elastic_f64.argtypes = [ctypes.POINTER(_EArgs), dp, dp, dp]
        L.oracle_ssg_deriv_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ip, dp]
        L.oracle_elastic_stagger_f32.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp, fp, ip, fp]
        L.oracle_elastic_stagger_f64.argtypes = [ctypes.c_int, ctypes.c_int, fp, fp, fp, ip, dp]
        L.oracle_elastic_smooth_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, fp]
        L._elastic_bound = True
    return L


def ssg_deriv(u, axis, flag, itopo=None):
    """One calux ("x") / caluz ("z") stencil sum of the elastic restatement
    (FP64) with flag = xflag / zflag; rows above itopo are left 0."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nx, nz = u.shape
    it = np.zeros(nx, np.int32) if itopo is None else np.ascontiguousarray(itopo, dtype=np.int32)
    out = np.zeros_like(u)
    rc = _elib().oracle_ssg_deriv_f64({"x": 0, "z": 1}[axis], flag, nx, nz, _ptr(u, ctypes.c_double),
                                      _ptr(it, ctypes.c_int), _ptr(out, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle_ssg_deriv_f64 returned {rc}")
    return out


def elastic_stagger(vp, vs, rho, itopo, precision="f32"):
    """substagger's staggered fields as a dict of [nx, nz] arrays."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    vp, vs, rho = f32(vp), f32(vs), f32(rho)
    nx, nz = vp.shape
    it = np.ascontiguousarray(itopo, dtype=np.int32)
    if precision == "f32":
        out = np.zeros((5, nx, nz), np.float32)
        rc = _elib().oracle_elastic_stagger_f32(nx, nz, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(it, ctypes.c_int),
                                                _ptr(out))
    else:
        out = np.zeros((5, nx, nz), np.float64)
        rc = _elib().oracle_elastic_stagger_f64(nx, nz, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(it, ctypes.c_int),
                                                _ptr(out, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle_elastic_stagger_f32 returned {rc}")
    return dict(zip(("busg2e", "busg1e", "lambmusg", "lambsg", "musge"), out))


def elastic_smooth(u, itopo, rgau):
    """modipapml's in-place Gaussian smoothing of rows 0..itopo (FP32)."""
    u = np.array(u, dtype=np.float32, order="C")
    nx, nz = u.shape
    it = np.ascontiguousarray(itopo, dtype=np.int32)
    rc = _elib().oracle_elastic_smooth_f32(nx, nz, rgau, _ptr(it, ctypes.c_int), _ptr(u))
    if rc != 0:
        raise ValueError(f"oracle_elastic_smooth_f32 returned {rc}")
    return u


def elastic(vp, vs, rho, itopo, *, src, rec, nsteps, delay=0, dx=5.0, dz=5.0, dt=1e-3, f0=25.0,
            sponge=(40, 40, 40), sponge_r=1e-3, rgau=50, wavelet=None, snap_steps=(), precision="f32",
            nthreads=1, mode="fsbc", npzn=0):
    """Run the elastic oracle (oracle/elastic_body.h), mode "fsbc" (free
    surface) or "abc" (absorbing top of npzn + 1 rows above the surface).
    Fields [nx, nz] (z fastest); itopo [nx] 0-based surface rows; sponge =
    (npxn, npxp, npzp).  Returns dict(rec_vx, rec_vz [nrec, nsteps - delay],
    snaps [nsnap, 5, nx, nz] = vx, vz, sxx, szz, sxz)."""
    L = _elib()
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    vp, vs, rho = f32(vp), f32(vs), f32(rho)
    nx, nz = vp.shape
    it = np.ascontiguousarray(itopo, dtype=np.int32)
    rx = np.ascontiguousarray(rec[0], dtype=np.int32)
    rz = np.ascontiguousarray(rec[1], dtype=np.int32)
    steps = np.ascontiguousarray(sorted(snap_steps), dtype=np.int32)
    w = f32(wavelet)
    if w is not None:
        assert w.shape[0] >= nsteps
    args = _EArgs(nx, nz, nsteps, delay, dx, dz, dt, f0, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(it, ctypes.c_int),
                  *sponge, sponge_r, int(src[0]), int(src[1]), rgau, _ptr(w), rx.size, _ptr(rx, ctypes.c_int),
                  _ptr(rz, ctypes.c_int), steps.size, _ptr(steps, ctypes.c_int), nthreads,
                  {"fsbc": 0, "abc": 1}[mode], npzn)
    dtype, ct = (np.float32, ctypes.c_float) if precision == "f32" else (np.float64, ctypes.c_double)
    nrt = max(nsteps - delay, 0)
    rvx = np.zeros((rx.size, nrt), dtype=dtype)
    rvz = np.zeros((rx.size, nrt), dtype=dtype)
    snaps = np.zeros((steps.size, 5, nx, nz), dtype=dtype)
    fn = L.oracle_elastic_f32 if precision == "f32" else L.oracle_elastic_f64
    rc = fn(ctypes.byref(args), _ptr(rvx, ct), _ptr(rvz, ct), _ptr(snaps, ct) if steps.size else ctypes.POINTER(ct)())
    if rc != 0:
        raise ValueError(f"oracle_elastic returned {rc}")
    return {"rec_vx": rvx, "rec_vz": rvz, "snaps": snaps}


# ---------------------------------------------------------------------- VTI
class _VTArgs(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nsteps", ctypes.c_int), ("delay", ctypes.c_int),
        ("nt", ctypes.c_int), ("dx", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("c2", ctypes.POINTER(ctypes.c_float)), ("c1", ctypes.POINTER(ctypes.c_float)),
        ("nl", ctypes.c_int), ("nr", ctypes.c_int), ("nu", ctypes.c_int), ("nd", ctypes.c_int),
        ("r", ctypes.c_float),
        ("vz", ctypes.POINTER(ctypes.c_float)), ("eps", ctypes.POINTER(ctypes.c_float)),
        ("dl", ctypes.POINTER(ctypes.c_float)),
        ("wavelet", ctypes.POINTER(ctypes.c_float)), ("nwav", ctypes.c_int),
        ("sx", ctypes.c_int), ("sz", ctypes.c_int),
        ("nrec", ctypes.c_int), ("rx", ctypes.POINTER(ctypes.c_int)), ("rz", ctypes.POINTER(ctypes.c_int)),
        ("nthreads", ctypes.c_int),
    ]


_vt = None


def _vtlib():
    global _vt
    if _vt is None:
        L = lib()
        fp, dp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.oracle_vti_f32.argtypes = [ctypes.POINTER(_VTArgs), fp, fp]
        L.oracle_vti_f64.argtypes = [ctypes.POINTER(_VTArgs), dp, dp]
        L.oracle_vti_coefficients_f32.argtypes = [ctypes.c_int, ctypes.c_int, fp]
        L.oracle_vti_wavelet_f32.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, fp]
        L.oracle_vti_window.argtypes = [fp, fp, fp] + [ctypes.c_int] * 9 + [fp, fp, fp]
        _vt = L
    return _vt


def vti_coefficients(kind: int, k: int = 5) -> np.ndarray:
    """coefficient_2nd (kind 2) / coefficient_1st (kind 1), FP32 as the reference."""
    out = np.zeros(k, dtype=np.float32)
    if _vtlib().oracle_vti_coefficients_f32(kind, k, _ptr(out)) != 0:
        raise ValueError("bad coefficient request")
    return out


def vti_wavelet(dt: float, f0: float) -> np.ndarray:
    """The VTI driver's wavelet (nwt samples, it = 1..nwt)."""
    L = _vtlib()
    n = L.oracle_vti_wavelet_f32(dt, f0, 0, ctypes.POINTER(ctypes.c_float)())
    w = np.zeros(n, dtype=np.float32)
    L.oracle_vti_wavelet_f32(dt, f0, n, _ptr(w))
    return w


def vti_window(vz, eps, delta, xmin, rng, pml, nsz0):
    """get_currshot_parameters + the isotropic top rows; model [nxv, nzv]."""
    vz, eps, delta = (np.ascontiguousarray(a, dtype=np.float32) for a in (vz, eps, delta))
    nxv, nzv = vz.shape
    nl, nr, nu, nd = pml
    shape = (rng + nl + nr - 2, nzv + nu + nd - 2)
    out = [np.zeros(shape, dtype=np.float32) for _ in range(3)]
    rc = _vtlib().oracle_vti_window(_ptr(vz), _ptr(eps), _ptr(delta), nxv, nzv, xmin, rng, nl, nr, nu, nd, nsz0,
                                    *(_ptr(o) for o in out))
    if rc != 0:
        raise ValueError("bad window")
    return tuple(out)


def vti(vz, eps, delta, *, src, rec, nsteps, delay=0, dx=5.0, dz=5.0, dt=5e-4, pml=(0, 0, 0, 0), r=1e-3,
        wavelet=None, c2=None, c1=None, precision="f32", nthreads=1, state=False):
    """Run the VTI oracle (oracle/vti_body.h).  Fields [nx, nz] (z fastest);
    pml = (left, right, top, bottom) widths; wavelet samples are added to p
    at src before steps 0, 1, ...  Returns the record [nrec, nsteps - delay]
    (and with state=True also the final p, q)."""
    L = _vtlib()
    vz, eps, delta = (np.ascontiguousarray(a, dtype=np.float32) for a in (vz, eps, delta))
    nx, nz = vz.shape
    c2 = vti_coefficients(2) if c2 is None else np.ascontiguousarray(c2, dtype=np.float32)
    c1 = vti_coefficients(1) if c1 is None else np.ascontiguousarray(c1, dtype=np.float32)
    w = np.zeros(0, dtype=np.float32) if wavelet is None else np.ascontiguousarray(wavelet, dtype=np.float32)
    rx = np.ascontiguousarray(rec[0], dtype=np.int32)
    rz = np.ascontiguousarray(rec[1], dtype=np.int32)
    nt = max(nsteps - delay, 0)
    args = _VTArgs(nx, nz, nsteps, delay, nt, dx, dz, dt, _ptr(c2), _ptr(c1), *pml, r, _ptr(vz), _ptr(eps),
                   _ptr(delta), _ptr(w), w.size, int(src[0]), int(src[1]), rx.size, _ptr(rx, ctypes.c_int),
                   _ptr(rz, ctypes.c_int), nthreads)
    dtype, ct = (np.float32, ctypes.c_float) if precision == "f32" else (np.float64, ctypes.c_double)
    record = np.zeros((rx.size, nt), dtype=dtype)
    st = np.zeros((2, nx, nz), dtype=dtype)
    fn = L.oracle_vti_f32 if precision == "f32" else L.oracle_vti_f64
    rc = fn(ctypes.byref(args), _ptr(record, ct), _ptr(st, ct))
    if rc != 0:
        raise ValueError(f"oracle_vti returned {rc}")
    return (record, st[0], st[1]) if state else record
