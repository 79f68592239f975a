"""C ABI checks that need no GPU: libseisfd.so loads, exports every function
declared in include/seisfd.h, and its host-only entry points answer."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"\b(sfd_\w+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import seismic_code_github_amd as sfd
    L = ctypes.CDLL(sfd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_covers_the_header():
    import seismic_code_github_amd as sfd
    assert set(declared_functions()) <= set(sfd._SIGNATURES), set(declared_functions()) - set(sfd._SIGNATURES)


def test_host_only_entry_points():
    import seismic_code_github_amd as sfd
    L = sfd.lib()
    assert L.sfd_abi_version() == 2
    # 1 / sum |C_m| for the reference's C1..C8 (Acoustic_1order_FD.cpp:38-45).
    c8 = np.array([1.2340911, -1.0664985e-1, 2.3036367e-2, -5.3423856e-3,
                   1.0772712e-3, -1.6641888e-4, 1.7021711e-5, -8.5234642e-7], dtype=np.float32)
    assert abs(sfd.cfl_limit(8) - 1.0 / np.abs(c8.astype(np.float64)).sum()) < 1e-12
    assert abs(sfd.cfl_limit(4) - 1.0 / (1225 / 1024 + 245 / 3072 + 49 / 5120 + 5 / 7168)) < 1e-6
    assert sfd.cfl_limit(3) == 0.0  # no default table for R = 3
    assert abs(sfd.cfl_limit(2, np.array([9 / 8, -1 / 24])) - 1 / (9 / 8 + 1 / 24)) < 1e-6


def test_params_struct_layout():
    import seismic_code_github_amd as sfd
    # int nx,nz,nt; float dx,dz,dt; int radius; ptr coef; int boundary,width;
    # float pml_r,pml_f0; int sx,sz,nrec; ptr rx,rz; int nsnap; ptr steps;
    # int snap_fields, arith
    assert ctypes.sizeof(sfd.AcousticParamsC) == 112  # sizeof(sfd_acoustic_params), x86-64


def test_elastic_params_struct_layout():
    import seismic_code_github_amd as sfd
    C = sfd.ElasticParamsC
    # sizeof / offsetof of sfd_elastic_params as gcc lays it out on x86-64
    assert ctypes.sizeof(C) == 120
    assert (C.itopo.offset, C.src_radius.offset, C.snap_steps.offset, C.snap_fields.offset) == (32, 64, 96, 104)


def test_every_struct_mirror_matches_the_header(tmp_path):
    """gcc compiles a probe against include/seisfd.h printing sizeof and the
    offset of every field the ctypes mirrors declare."""
    import subprocess
    import seismic_code_github_amd as sfd
    mirrors = {"sfd_acoustic_params": sfd.AcousticParamsC, "sfd_visco_params": sfd.ViscoParamsC,
               "sfd_elastic_params": sfd.ElasticParamsC, "sfd_vti_params": sfd.VtiParamsC,
               "sfd_srm_params": sfd.SrmParamsC,
               "sfd_q_par": sfd.QParC, "sfd_iso_par": sfd.IsoParC}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "seisfd.h"', 'int main(void) {']
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ['return 0;', '}']
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, cls in mirrors.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_params_arrays_stay_alive_for_repeated_shots():
    """The same shot object twice in one batch or gradient call: the struct of
    the first entry must still point at live receiver arrays after the second
    struct is made (each struct's arrays go to the caller's keep list)."""
    import gc
    import seismic_code_github_amd as sfd
    from seismic_code_github_amd import workloads as W
    shot = W.layered_301(nt=10)
    keep = []
    params = (sfd.AcousticParamsC * 2)()
    for i in range(2):
        params[i] = shot.params(keep)
    gc.collect()
    junk = [np.full(301, -7, dtype=np.int32) for _ in range(64)]  # reuse freed memory, if any
    rx = np.asarray(shot.rec[0], dtype=np.int32)
    rz = np.asarray(shot.rec[1], dtype=np.int32)
    for i in range(2):
        got_x = np.ctypeslib.as_array(params[i].rx, shape=(params[i].nrec,))
        got_z = np.ctypeslib.as_array(params[i].rz, shape=(params[i].nrec,))
        np.testing.assert_array_equal(got_x, rx)
        np.testing.assert_array_equal(got_z, rz)
    assert len(keep) == 8 and len(junk) == 64
