"""CPU-side checks of the product: the C-ABI library loads and exports every symbol the header
declares, host-side set-up matches the oracle, the Fortran ISO_C_BINDING layer compiles and
links, and the parameter plumbing round-trips.  No device calls."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from rpmd_mainz_amd import configs as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "rpmd_mi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rpmd_\w+)\s*\(", hdr, re.M)))


def test_library_exports_header_symbols():
    from rpmd_mainz_amd import capi
    L = capi.lib()
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s


def test_setup_ewald_matches_oracle(oracle):
    from oracle.oracle import RpmdParams
    from rpmd_mainz_amd import capi
    for ncell, aspect, ecut in [((12, 12, 12), (1, 1, 1), 1e-6), ((6, 6, 6), (1, 1, 1), 0.0),
                                ((16, 16, 16), (1, 1, 1.2), 1e-6), ((5, 5, 5), (1, 1, 1), 0.0)]:
        r, box, z, mass, prm = cf.water_system("qspcfw", ncell, 1, aspect=aspect, ecut=ecut)
        na = r.shape[1]
        a = capi.setup_ewald(na, box, dict(prm, iamcub=1))
        P = RpmdParams.from_dict(dict(prm, iamcub=1))
        oracle.setup_ewald(na, box, P)
        for k in ("wrcut", "walpha", "rkmax", "kmax", "iamcub"):
            assert a[k] == getattr(P, k), (ncell, k)
            assert prm[k] == getattr(P, k), (ncell, k)


def test_config_kmax_of_the_benchmark_configs():
    # the survey's derived Ewald parameters (unit box: wrcut, kmax)
    r, box, z, mass, prm = cf.water_system("qspcfw", (12, 12, 12), 1, ecut=1e-6)
    assert prm["kmax"] == 15 and abs(prm["wrcut"] * box[0] * cf.TOA - 10.76) < 0.01
    r, box, z, mass, prm = cf.water_system("qspcfw", (6, 6, 6), 1, ecut=1e-6)
    assert prm["kmax"] == 10


def test_params_struct_layout_matches_header():
    import ctypes as C
    from oracle.oracle import RpmdParams as P1
    from rpmd_mainz_amd.capi import RpmdParams as P2
    assert C.sizeof(P1) == C.sizeof(P2)
    assert [f[0] for f in P1._fields_] == [f[0] for f in P2._fields_]
    hdr = open(os.path.join(ROOT, "include", "rpmd_mi355x.h")).read()
    body = hdr[hdr.index("typedef struct rpmd_params"):hdr.index("} rpmd_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(\w+)\s*[,;]", body)
    assert names == [f[0] for f in P2._fields_]


@pytest.mark.skipif(shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"),
                    reason="amdflang absent")
def test_fortran_shim_compiles_and_links(tmp_path):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    src = os.path.join(ROOT, "rpmd_mainz_amd", "fortran", "rpmd_hip.f90")
    obj = tmp_path / "rpmd_hip.o"
    subprocess.run([fc, "-c", src, "-J", str(tmp_path), "-o", str(obj)], check=True)
    # a caller written against the reference's implicit interfaces
    caller = tmp_path / "caller.f90"
    caller.write_text("program c\n  real(8) r(3,3,2), d(3,3,2), z(3), b(3), v, w(3,3)\n"
                      "  call forces(r, v, d, 2, 3, b, z, w, 0)\n  call realft(r, 9, 2, 1)\nend program c\n")
    exe = tmp_path / "caller"
    subprocess.run([fc, str(caller), str(obj), "-L", os.path.join(ROOT, "rpmd_mainz_amd"), "-lrpmd_mi355x",
                    "-o", str(exe)], check=True)
    assert exe.exists()


def test_no_cpu_fallback_in_product():
    """The product path never loads the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "rpmd_mainz_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".f90")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower() or f == "configs.py" and "oracle" not in txt, f


class _PairTab(__import__("ctypes").Structure):
    import ctypes as _c
    _fields_ = [("d", _c.c_void_p), ("c", _c.c_float), ("ku", _c.c_int), ("kr", _c.c_int),
                ("key", _c.c_double * 3)]


def _as_pair(s, a, sfac):
    """du = erfc_AS(a r)/r, dur = (sfac e^{-a^2 s} - du)/s (A&S 7.1.26, the reference's
    src/ewald_cubic.f90:126-140), in FP64"""
    p, A = 0.3275911, (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)
    r = np.sqrt(s)
    x = a * r
    e = np.exp(-x * x)
    t = 1.0 / (1.0 + p * x)
    h = e * t * (A[0] + t * (A[1] + t * (A[2] + t * (A[3] + t * A[4]))))
    du = h / r
    return du, (sfac * e - du) / s


@pytest.mark.parametrize("alpha,rc,sr", [(12.885849527615372, 0.2884499140614817, 2.0 ** -32),
                                         (11.809706162344863, 0.31473451902649446, 2.0 ** -32),
                                         (0.17994, 17.4593, 1.0)])
def test_fp32_pair_table_against_as_form(alpha, rc, sr):
    """The FP32 real-space pair-function table (host builder of the library), looked up exactly as
    the kernel does (FP32 s' = c r^2 clamped to [2, 2^14], interval from the low 4 exponent and top 5
    mantissa bits, cubic in the low 18 mantissa bits): du and dur within 1e-5 of the A&S form
    (written here with the textbook t = 1 / (1 + 0.3275911 x), which the reference's
    t = 3.0525860 / (3.0525860 + x) matches to 2e-7), zero at and past the cut-off, the cut-off at
    s' = 2^14."""
    import ctypes
    from rpmd_mainz_amd import capi
    L = capi.lib()
    fill = getattr(L, "_ZN4rpmd15pair_table_fillEdddRNS_7PairTabEP15HIP_vector_typeIfLj4EE")
    fill.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_PairTab), ctypes.c_void_p]
    n = 13 * 32 + 1
    h = np.zeros((2, n, 4), np.float32)
    t = _PairTab()
    fill(alpha, rc, sr, ctypes.byref(t), h.ctypes.data)
    s_rc = (rc / sr) ** 2
    assert abs(float(np.float32(t.c)) * s_rc / 2.0 ** 14 - 1.0) < 1e-7
    assert np.all(h[:, n - 1] == 0.0)
    a, sfac = alpha * sr, -2.0 * alpha / np.sqrt(np.pi) * sr
    rng = np.random.default_rng(1)
    s = np.exp(rng.uniform(np.log(s_rc * 2.0 ** -12.9), np.log(s_rc), 200000))
    r2 = s.astype(np.float32)
    sp = np.clip(np.float32(t.c) * r2, np.float32(2.0), np.float32(2.0 ** 14))
    bits = sp.view(np.uint32)
    idx = (bits >> np.uint32(18)) & np.uint32(0x1FF)
    w = (bits & np.uint32(0x3FFFF)).astype(np.float32)
    ev = lambda c: ((c[:, 3] * w + c[:, 2]) * w + c[:, 1]) * w + c[:, 0]
    du = ev(h[0, idx]).astype(np.float64) * 2.0 ** -t.ku
    dur = ev(h[1, idx]).astype(np.float64) * 2.0 ** -t.kr
    du_x, dur_x = _as_pair(r2.astype(np.float64), a, sfac)
    inside = np.float32(t.c) * r2 < np.float32(2.0 ** 14)
    assert np.abs(du[inside] / du_x[inside] - 1.0).max() < 1e-5
    assert np.abs(dur[inside] / dur_x[inside] - 1.0).max() < 1e-5
    assert np.all(du[~inside] == 0.0) and np.all(dur[~inside] == 0.0)
