"""Python mirror of the reference's force-routine interface over the HIP C-ABI.

The entry points keep the reference's names, argument meaning and error behaviour
(src/md_forces.f90:1 ``forces``, :305 ``full_forces``; src/fourier.f90:16 ``realft``;
src/evolve_pi.f90:1 ``evolve_pi``).  Arrays are numpy float64 in the reference's memory layout:
a Fortran ``r(3,na,nb)`` is a C-contiguous ``(nb, na, 3)`` array here.  A STOP in the reference
is an ``RPMDError`` carrying the reference's message.

Everything runs on the GPU through ``librpmd_mi355x.so``; there is no CPU fallback -- loading
fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "librpmd_mi355x.so")

PREC = {"fp64": 0, "mixed": 1}

_dp = C.POINTER(C.c_double)


class RpmdParams(C.Structure):
    """``rpmd_params`` of include/rpmd_mi355x.h (the reference COMMON blocks)."""

    _fields_ = [
        ("oo_eps", C.c_double), ("oo_sig", C.c_double), ("oo_gam", C.c_double), ("rcut", C.c_double),
        ("alpha", C.c_double), ("ecut", C.c_double), ("wm", C.c_double), ("wh", C.c_double),
        ("wrcut", C.c_double), ("walpha", C.c_double), ("rkmax", C.c_double),
        ("kmax", C.c_int), ("iamcub", C.c_int),
        ("apot", C.c_double), ("bpot", C.c_double), ("alp", C.c_double), ("alpb", C.c_double),
        ("theta", C.c_double), ("reoh", C.c_double),
        ("sig", C.c_double),
        ("nbdf1", C.c_int), ("nbdf2", C.c_int), ("mts", C.c_int),
        ("intra_model", C.c_int), ("therm", C.c_int), ("ttau", C.c_double),
        ("rigid", C.c_int), ("rat_tol", C.c_double),
    ]

    @classmethod
    def from_dict(cls, d: dict) -> "RpmdParams":
        p = cls()
        for k, v in d.items():
            setattr(p, k, v)
        return p


class RPMDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (rpmd error {code})")
        self.code = code


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIBPATH)
        L.rpmd_strerror.restype = C.c_char_p
        for name in ("rpmd_setup_ewald", "rpmd_ctx_create", "rpmd_ctx_set_params", "rpmd_ctx_set_precision",
                     "rpmd_ctx_destroy", "rpmd_forces", "rpmd_full_forces", "rpmd_realft", "rpmd_evolve_pi",
                     "rpmd_state_upload", "rpmd_run_steps", "rpmd_state_download", "rpmd_use_graph",
                     "rpmd_prof_begin", "rpmd_prof_end", "rpmd_gle_init", "rpmd_evolve_rig_pi",
                     "rpmd_batch_create", "rpmd_batch_run_steps", "rpmd_batch_destroy", "rpmd_sample_rpmd",
                     "rpmd_ctx_set_streams", "rpmd_free_ring", "rpmd_realft_noalloc", "rpmd_estimators",
                     "rpmd_batch_tuning", "rpmd_ipi_run", "rpmd_set_barostat", "rpmd_barostat_counts",
                     "rpmd_evolve_pi_npt", "rpmd_evolve_rig_pi_npt"):
            getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def _check(code: int):
    if code != 0:
        raise RPMDError(code, lib().rpmd_strerror(code).decode())


def _arr(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def run_driver(ctx: "RPMDContext", na: int, nb: int, z, serverdata: str = "serverdata") -> int:
    """subroutine run_driver (src/ipi_driver.f90:1-254): nb must be 1; the server address is the first
    whitespace-separated word of the file serverdata, as the reference reads it (list-directed)."""
    if nb != 1:
        raise ValueError("i-PI mode must have nb = 1!")
    with open(serverdata) as f:
        address = f.read().split()[0].strip("'\"")
    return ctx.ipi_run(address, z, na)


def setup_ewald(na: int, boxlxyz, prm: dict) -> dict:
    """setup_ewald (src/setup_options.f90:104-189) through the C-ABI."""
    P = RpmdParams.from_dict(prm)
    b = _arr(boxlxyz, (3,))
    _check(lib().rpmd_setup_ewald(C.c_int(na), _p(b), C.byref(P)))
    out = dict(prm)
    for k in ("wrcut", "walpha", "rkmax", "kmax", "iamcub"):
        out[k] = getattr(P, k)
    return out


def realft_noalloc(data: np.ndarray, m: int, n: int, mode: int) -> np.ndarray:
    """rpmd_realft_noalloc: realft (src/fourier.f90:16) without a context and without any device
    allocation, in place on a C-contiguous (n, m) array (= Fortran data(m,n)); also returned."""
    if not (isinstance(data, np.ndarray) and data.dtype == np.float64 and data.flags.c_contiguous):
        raise TypeError("realft_noalloc needs a C-contiguous float64 array")
    _check(lib().rpmd_realft_noalloc(_p(data), C.c_int(m), C.c_int(n), C.c_int(mode)))
    return data


class RPMDContext:
    """One device context (buffers for na atoms x nb_max beads) with the reference's entry points."""

    def __init__(self, prm: dict, na: int, nb_max: int, precision: str = "fp64"):
        self._P = RpmdParams.from_dict(prm)
        self.prm = dict(prm)
        self.na, self.nb_max = na, nb_max
        h = C.c_void_p()
        _check(lib().rpmd_ctx_create(C.byref(self._P), C.c_int(na), C.c_int(nb_max), C.c_int(PREC[precision]),
                                     C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().rpmd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, prm: dict):
        self._P = RpmdParams.from_dict(prm)
        self.prm = dict(prm)
        _check(lib().rpmd_ctx_set_params(self._h, C.byref(self._P)))

    def set_precision(self, precision: str):
        _check(lib().rpmd_ctx_set_precision(self._h, C.c_int(PREC[precision])))

    # -- subroutine forces(r,v,dvdr,nb,na,boxlxyz,z,virial,iopt)  src/md_forces.f90:1 ----------
    def forces(self, r, nb: int, na: int, boxlxyz, z, iopt: int):
        r = _arr(r, (nb, na, 3))
        dvdr = np.zeros_like(r)
        vir = np.zeros(9)
        v = C.c_double()
        _check(lib().rpmd_forces(self._h, _p(r), C.c_int(nb), C.c_int(na), _p(_arr(boxlxyz, (3,))),
                                 _p(_arr(z, (na,))), C.c_int(iopt), C.byref(v), _p(dvdr), _p(vir)))
        return v.value, dvdr, vir.reshape(3, 3).T.copy()

    # -- subroutine full_forces(r,na,nb,v,vew,voo,vint,vir,z,boxlxyz,dvdr,dvdr2) :305 ---------
    def full_forces(self, r, na: int, nb: int, z, boxlxyz):
        r = _arr(r, (nb, na, 3))
        dvdr, dvdr2 = np.zeros_like(r), np.zeros_like(r)
        vir = np.zeros(9)
        v, vew, voo, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check(lib().rpmd_full_forces(self._h, _p(r), C.c_int(na), C.c_int(nb), C.byref(v), C.byref(vew),
                                      C.byref(voo), C.byref(vint), _p(vir), _p(_arr(z, (na,))),
                                      _p(_arr(boxlxyz, (3,))), _p(dvdr), _p(dvdr2)))
        return dict(v=v.value, vew=vew.value, voo=voo.value, vint=vint.value, vir=vir.reshape(3, 3).T.copy(),
                    dvdr=dvdr, dvdr2=dvdr2)

    # -- subroutine run_driver(...)  src/ipi_driver.f90:1, src/sockets.c ---------------------------
    def ipi_run(self, address: str, z, na: int) -> int:
        """Serve the i-PI server at address ("unix:<name>:<port>" for the UNIX socket /tmp/ipi_<name>,
        "<any>:<host>:<port>" for an internet socket) with full_forces on one bead until it disconnects
        or sends EXIT.  Returns the number of force replies sent."""
        served = C.c_int()
        _check(lib().rpmd_ipi_run(self._h, address.encode(), _p(_arr(z, (na,))), C.c_int(na), C.byref(served)))
        return served.value

    # -- subroutine realft(data,m,n,mode)  src/fourier.f90:16 --------------------------------------
    def realft(self, data: np.ndarray, m: int, n: int, mode: int) -> np.ndarray:
        """In place on a C-contiguous (n, m) array (= Fortran data(m,n)); also returned."""
        if not (isinstance(data, np.ndarray) and data.dtype == np.float64 and data.flags.c_contiguous):
            raise TypeError("realft needs a C-contiguous float64 array")
        _check(lib().rpmd_realft(self._h, _p(data), C.c_int(m), C.c_int(n), C.c_int(mode)))
        return data

    # -- free ring-polymer update of evolve_pi  src/evolve_pi.f90:109-125 ----------------------
    def free_ring(self, p, r, mass, na: int, nb: int, dt: float, beta: float):
        """p, r: normal-mode coordinates (nb, na, 3), updated in place and returned."""
        for a in (p, r):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (nb, na, 3)):
                raise TypeError("free_ring arrays must be C-contiguous float64 (nb, na, 3)")
        _check(lib().rpmd_free_ring(self._h, _p(p), _p(r), _p(_arr(mass, (na,))), C.c_int(na), C.c_int(nb),
                                    C.c_double(dt), C.c_double(beta)))
        return p, r

    # -- subroutine evolve_pi(p,r,v,vew,vlj,vint,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,
    #                         vir_lf,irun,nbaro)  src/evolve_pi.f90:1 ------------------------------
    def evolve_pi(self, p, r, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, noise=None):
        """One step; p, r, dvdr, dvdr2 are updated in place (C-contiguous float64 (nb,na,3)).
        noise: Gaussian deviates in the reference draw order, or None for the device stream."""
        for a in (p, r, dvdr, dvdr2):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (nb, na, 3)):
                raise TypeError("evolve_pi state arrays must be C-contiguous float64 (nb, na, 3)")
        vir, vir_lf = np.zeros(9), np.zeros(9)
        v, vew, vlj, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        if noise is not None:
            noise = _arr(noise)
            npt, nn = _p(noise), C.c_longlong(noise.size)
        else:
            npt, nn = None, C.c_longlong(0)
        _check(lib().rpmd_evolve_pi(self._h, _p(p), _p(r), C.byref(v), C.byref(vew), C.byref(vlj), C.byref(vint),
                                    _p(dvdr), _p(dvdr2), C.c_double(dt), _p(_arr(mass, (na,))), C.c_int(na),
                                    C.c_int(nb), _p(_arr(boxlxyz, (3,))), _p(_arr(z, (na,))), C.c_double(beta),
                                    _p(vir), _p(vir_lf), npt, nn))
        return dict(v=v.value, vew=vew.value, vlj=vlj.value, vint=vint.value, vir=vir.reshape(3, 3).T.copy(),
                    vir_lf=vir_lf.reshape(3, 3).T.copy())

    # -- module barostat (src/module_barostat.f90:1-5) and evolve_pi with nbaro = 1 ----------------
    BARO = {"NONE": 0, "BER": 1, "MCI": 2}

    def set_barostat(self, baro: str, taub: float, patm: float, dv: float = 200.0):
        """baro 'BER' (taub = 41.341373 taufs) or 'MCI' (taub = taufs, dv in bohr^3); patm in a.u."""
        if baro not in self.BARO:
            raise RPMDError(21, " ** Invalid barostat **")
        _check(lib().rpmd_set_barostat(self._h, C.c_int(self.BARO[baro]), C.c_double(taub), C.c_double(patm),
                                       C.c_double(dv)))

    def barostat_counts(self):
        """(nacc, nmove) of the MC barostat"""
        a, m = C.c_longlong(), C.c_longlong()
        _check(lib().rpmd_barostat_counts(self._h, C.byref(a), C.byref(m)))
        return a.value, m.value

    def evolve_pi_npt(self, p, r, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir_lf, noise=None,
                      uniforms=None):
        """One evolve_pi step with nbaro = 1; p, r, dvdr, dvdr2 updated in place; vir_lf (3x3) the previous
        step's low-frequency virial; uniforms: the MC barostat's ran2 draws in order, or None for the
        library's stream.  Returns energies, virials and the new box."""
        for a in (p, r, dvdr, dvdr2):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (nb, na, 3)):
                raise TypeError("evolve_pi state arrays must be C-contiguous float64 (nb, na, 3)")
        vir = np.zeros(9)
        vl = np.ascontiguousarray(np.asarray(vir_lf, dtype=np.float64).T).reshape(9).copy()
        box = _arr(boxlxyz, (3,)).copy()
        v, vew, vlj, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        if noise is not None:
            noise = _arr(noise)
            npt, nn = _p(noise), C.c_longlong(noise.size)
        else:
            npt, nn = None, C.c_longlong(0)
        if uniforms is not None:
            uniforms = _arr(uniforms)
            up, nu = _p(uniforms), C.c_int(uniforms.size)
        else:
            up, nu = None, C.c_int(0)
        _check(lib().rpmd_evolve_pi_npt(self._h, _p(p), _p(r), C.byref(v), C.byref(vew), C.byref(vlj), C.byref(vint),
                                        _p(dvdr), _p(dvdr2), C.c_double(dt), _p(_arr(mass, (na,))), C.c_int(na),
                                        C.c_int(nb), _p(box), _p(_arr(z, (na,))), C.c_double(beta), _p(vir), _p(vl),
                                        npt, nn, up, nu))
        return dict(v=v.value, vew=vew.value, vlj=vlj.value, vint=vint.value, vir=vir.reshape(3, 3).T.copy(),
                    vir_lf=vl.reshape(3, 3).T.copy(), box=box)

    # -- subroutine evolve_rig_pi(p,r,v,v_lf,v_hf,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,
    #                             vir_lf,irun,nbaro)  src/evolve_rigid.f90:100 ---------------------
    def evolve_rig_pi(self, p, r, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir=None, noise=None):
        """One rigid step; p, r, dvdr, dvdr2 updated in place; vir (3x3) in/out as in the reference."""
        for a in (p, r, dvdr, dvdr2):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (nb, na, 3)):
                raise TypeError("evolve_rig_pi state arrays must be C-contiguous float64 (nb, na, 3)")
        w = np.zeros(9) if vir is None else np.ascontiguousarray(np.asarray(vir, dtype=np.float64).T).reshape(9)
        w = w.copy()
        wl = np.zeros(9)
        v, vlf, vhf = C.c_double(), C.c_double(), C.c_double()
        if noise is not None:
            noise = _arr(noise)
            npt, nn = _p(noise), C.c_longlong(noise.size)
        else:
            npt, nn = None, C.c_longlong(0)
        _check(lib().rpmd_evolve_rig_pi(self._h, _p(p), _p(r), C.byref(v), C.byref(vlf), C.byref(vhf), _p(dvdr),
                                        _p(dvdr2), C.c_double(dt), _p(_arr(mass, (na,))), C.c_int(na), C.c_int(nb),
                                        _p(_arr(boxlxyz, (3,))), _p(_arr(z, (na,))), C.c_double(beta), _p(w),
                                        _p(wl), npt, nn))
        return dict(v=v.value, v_lf=vlf.value, v_hf=vhf.value, vir=w.reshape(3, 3).T.copy(),
                    vir_lf=wl.reshape(3, 3).T.copy())

    def evolve_rig_pi_npt(self, p, r, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir=None, noise=None,
                          uniforms=None):
        """One rigid step with nbaro = 1 (src/evolve_rigid.f90:203-220); as evolve_rig_pi, and the new box
        is returned; uniforms: the MC barostat's ran2 draws in order, or None for the library's stream."""
        for a in (p, r, dvdr, dvdr2):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
                    and a.shape == (nb, na, 3)):
                raise TypeError("evolve_rig_pi state arrays must be C-contiguous float64 (nb, na, 3)")
        w = np.zeros(9) if vir is None else np.ascontiguousarray(np.asarray(vir, dtype=np.float64).T).reshape(9)
        w = w.copy()
        wl = np.zeros(9)
        box = _arr(boxlxyz, (3,)).copy()
        v, vlf, vhf = C.c_double(), C.c_double(), C.c_double()
        if noise is not None:
            noise = _arr(noise)
            npt, nn = _p(noise), C.c_longlong(noise.size)
        else:
            npt, nn = None, C.c_longlong(0)
        if uniforms is not None:
            uniforms = _arr(uniforms)
            up, nu = _p(uniforms), C.c_int(uniforms.size)
        else:
            up, nu = None, C.c_int(0)
        _check(lib().rpmd_evolve_rig_pi_npt(self._h, _p(p), _p(r), C.byref(v), C.byref(vlf), C.byref(vhf), _p(dvdr),
                                            _p(dvdr2), C.c_double(dt), _p(_arr(mass, (na,))), C.c_int(na), C.c_int(nb),
                                            _p(box), _p(_arr(z, (na,))), C.c_double(beta), _p(w), _p(wl), npt, nn, up,
                                            nu))
        return dict(v=v.value, v_lf=vlf.value, v_hf=vhf.value, vir=w.reshape(3, 3).T.copy(),
                    vir_lf=wl.reshape(3, 3).T.copy(), box=box)

    # -- subroutine sample_rpmd(p,na,nb,mass,beta,irun)  src/module_thermostat.f90:199 ----------
    def sample_rpmd(self, nb: int, na: int, mass, beta: float, noise=None, seed: int = 0) -> np.ndarray:
        """Bead momenta (nb, na, 3); noise: standard normals in the reference draw order, or None."""
        p = np.zeros((nb, na, 3))
        if noise is not None:
            noise = _arr(noise)
            npt, nn = _p(noise), C.c_longlong(noise.size)
        else:
            npt, nn = None, C.c_longlong(0)
        _check(lib().rpmd_sample_rpmd(self._h, _p(p), C.c_int(na), C.c_int(nb), _p(_arr(mass, (na,))),
                                      C.c_double(beta), npt, nn, C.c_ulonglong(seed)))
        return p

    # -- subroutine therm_gle_init(ttau,na,nb,dt,irun,beta)  src/module_gle.f90:146 ---------------
    def gle_init(self, A, na: int, nb: int, dt: float, beta: float, gp=None, seed: int = 0):
        """A: (ns+1, ns+1) drift matrix as the GLE-A file lists it (rows); dt: the half step the
        reference passes (src/main.f90:609); gp: initial auxiliary momenta, shape (ns+1, nb, na, 3)
        (= Fortran gp(3,na,nb,ns+1)), or None to draw them on the device."""
        A = _arr(A)
        ns = A.shape[0] - 1
        gpp = _p(_arr(gp, (ns + 1, nb, na, 3))) if gp is not None else None
        _check(lib().rpmd_gle_init(self._h, C.c_int(ns), _p(A), C.c_int(na), C.c_int(nb), C.c_double(dt),
                                   C.c_double(beta), gpp, C.c_ulonglong(seed)))

    # -- device-resident stepping (no reference counterpart) ---------------------------------------
    def state_upload(self, p, r, mass, z, na, nb, boxlxyz, beta, dt, seed=12345):
        _check(lib().rpmd_state_upload(self._h, _p(_arr(p, (nb, na, 3))), _p(_arr(r, (nb, na, 3))),
                                       _p(_arr(mass, (na,))), _p(_arr(z, (na,))), C.c_int(na), C.c_int(nb),
                                       _p(_arr(boxlxyz, (3,))), C.c_double(beta), C.c_double(dt),
                                       C.c_ulonglong(seed)))
        self._st_nb = nb

    def use_graph(self, enable: bool = True):
        _check(lib().rpmd_use_graph(self._h, C.c_int(1 if enable else 0)))

    def set_streams(self, streams: int):
        """2: the real-space Ewald branch on a second stream (default); 3: also the O-O and
        intramolecular terms on a third; 1: one stream"""
        _check(lib().rpmd_ctx_set_streams(self._h, C.c_int(streams)))

    def run_steps(self, nsteps: int):
        _check(lib().rpmd_run_steps(self._h, C.c_int(nsteps)))

    def prof_begin(self):
        """Start HIP-event timing of the Ewald kernels (eager launches on the context stream)."""
        _check(lib().rpmd_prof_begin(self._h))

    def prof_end(self) -> dict:
        ms = (C.c_double * 5)()
        cnt = (C.c_int * 5)()
        _check(lib().rpmd_prof_end(self._h, ms, cnt))
        names = ("ewald_real", "k_sk", "k_kforce", "k_sk_final", "vlist")
        return {n: (ms[i], cnt[i]) for i, n in enumerate(names)}

    EST_NAMES = ("kinetic", "springs", "potential", "h_rp_per_atom", "temperature_K", "tv", "tq1", "tq2",
                 "tv_x", "tv_y", "tv_z")

    def estimators(self) -> dict:
        """rpmd_estimators: kinetic and spring energies, H_RP per atom, temperature and the
        centroid-virial kinetic energy of the device-resident state, computed on the device."""
        out = (C.c_double * len(self.EST_NAMES))()
        _check(lib().rpmd_estimators(self._h, out))
        return {n: out[i] for i, n in enumerate(self.EST_NAMES)}

    def state_download(self):
        nb, na = self._st_nb, self.na
        p, r = np.zeros((nb, na, 3)), np.zeros((nb, na, 3))
        dvdr, dvdr2 = np.zeros((nb, na, 3)), np.zeros((nb, na, 3))
        e, vir = np.zeros(4), np.zeros(9)
        _check(lib().rpmd_state_download(self._h, _p(p), _p(r), _p(dvdr), _p(dvdr2), _p(e), _p(vir)))
        return dict(p=p, r=r, dvdr=dvdr, dvdr2=dvdr2, v=e[0], vew=e[1], vlj=e[2], vint=e[3],
                    vir=vir.reshape(3, 3).T.copy())


class ReplicaBatch:
    """Independent replicas (each an RPMDContext with an uploaded state) stepped by one hipGraph."""

    def __init__(self, contexts):
        self._ctx = list(contexts)
        arr = (C.c_void_p * len(self._ctx))(*[c._h.value for c in self._ctx])
        h = C.c_void_p()
        _check(lib().rpmd_batch_create(arr, C.c_int(len(self._ctx)), C.byref(h)))
        self._h = h

    def run_steps(self, nsteps: int):
        _check(lib().rpmd_batch_run_steps(self._h, C.c_int(nsteps)))

    def tuning(self) -> dict:
        """the autotuners' choices: batched steps in use (1), separate (0) or none yet (-1), the timed
        pair of steps of each form (ms), the pair walk's cap and whether it runs beside k space"""
        bt, cap, two = C.c_int(), C.c_int(), C.c_int()
        mf, ms = C.c_double(), C.c_double()
        _check(lib().rpmd_batch_tuning(self._h, C.byref(bt), C.byref(mf), C.byref(ms), C.byref(cap), C.byref(two)))
        return dict(batched=bt.value, ms_batched_2steps=mf.value, ms_separate_2steps=ms.value, pair_cap=cap.value,
                    pair_beside=two.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().rpmd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
