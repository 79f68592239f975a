"""Small calls of every oracle entry point, run in a subprocess against the
AddressSanitizer / UndefinedBehaviorSanitizer build of the oracle
(SEISFD_ORACLE_LIB; tests/test_sanitized.py).  Exit 0 = no report."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

assert oracle.LIB_PATH == os.environ["SEISFD_ORACLE_LIB"]
rng = np.random.default_rng(0)
nx, nz = 40, 34
vp = (1800 + 400 * rng.random((nx, nz))).astype(np.float32)
den = np.full((nx, nz), 2200.0, np.float32)
w = oracle.ricker(80, np.float32(0.001), 30.0, 50.0)
rec = (np.arange(5, 35, dtype=np.int32), np.full(30, 6, np.int32))
for precision in ("f32", "f64"):
    oracle.acoustic(vp, den, w, (20, 10), rec, dt=0.001, radius=4, sponge=8, snap_steps=[3, 79],
                    precision=precision, nthreads=2)
    oracle.acoustic(vp, den, w, (20, 10), rec, dt=0.001, radius=8, sponge=10, boundary="cpml", pml_f0=30.0,
                    precision=precision, nthreads=2)
oracle.gradient(vp, den, w, (20, 10), rec, np.ones((30, 80)), dt=0.001, radius=5, sponge=8)
oracle.window(vp, 3, 20, 4, 4, 3, 3)
q = np.full((nx, nz), 40.0, np.float32)
for nmech in (1, 3):
    oracle.visco(vp, den, q, src=(20, 10), rec=rec, nsteps=60, sponge=(5, 5, 5, 5), nmech=nmech,
                 vs=vp / 1.8, snap_steps=[59], precision="f64", nthreads=2)
init = np.zeros((5, nx, nz), np.float32)
init[4, 18:22] = 1.0
oracle.visco(vp, den, q, src=(0, 0), rec=rec, nsteps=20, sponge=(0, 0, 0, 0), vs=vp / 1.8, init=init,
             wavelet=np.zeros(20, np.float32))
oracle.rsg_deriv(rng.standard_normal((20, 18)), "z", 1)
oracle.csponge(12, 10.0)
itopo = np.full(nx, 2, np.int32)
for mode, npzn in (("fsbc", 0), ("abc", 2)):
    oracle.elastic(vp, vp / 1.7, den, itopo, src=(20, 6), rec=rec, nsteps=50, nthreads=2, sponge=(8, 8, 8),
                   rgau=6, mode=mode, npzn=npzn, snap_steps=[49], precision="f64")
oracle.elastic_stagger(vp, vp / 1.7, den, itopo)
oracle.ssg_deriv(rng.standard_normal((nx, nz)), "z", 0, itopo)
ep = np.full((nx, nz), 0.1, np.float32)
oracle.vti(vp, ep, ep * 0.5, src=(20, 10), rec=rec, nsteps=60, pml=(6, 6, 0, 6), precision="f64", nthreads=2)
oracle.vti_coefficients(2, 5)
print("oracle calls ok")
