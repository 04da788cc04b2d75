"""ctypes front-end to the CPU checkers (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker or the timed CPU baseline -- never as part of the product path.

* ``Oracle``  -- this repository's C restatement of the reference path (liboracle.so).
* ``RefPath`` -- the reference's own Fortran force path, RNG and thermostat helpers compiled in
  place from the reference sources by ``oracle/Makefile`` (``_ref/libref_path.so``).  The
  reference COMMON blocks are set through their exported symbols and the Fortran routines are
  called with by-reference arguments, exactly as the Fortran callers would.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libref_path.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class RpmdParams(C.Structure):
    """Mirror of ``rpmd_params`` in include/rpmd_mi355x.h."""

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

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class RngState(C.Structure):
    _fields_ = [("idum", C.c_int), ("idum2", C.c_int), ("iy", C.c_int), ("iv", C.c_int * 32),
                ("iset", C.c_int), ("gset", C.c_double)]


class GleState(C.Structure):
    _fields_ = [("ns", C.c_int), ("na", C.c_int), ("nb", C.c_int),
                ("gT", C.c_double * 64), ("gS", C.c_double * 64), ("gp", _dp)]


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The C restatement (oracle/rpmd_oracle.c)."""

    def __init__(self, path: str = LIB_ORACLE):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_forces.restype = C.c_int
        L.orc_full_forces.restype = C.c_int
        L.orc_realft.restype = C.c_int
        L.orc_evolve_pi.restype = C.c_int
        L.orc_evolve_pi_npt.restype = C.c_int
        L.orc_evolve_rig_pi_npt.restype = C.c_int
        L.orc_mc_baro.restype = C.c_int
        L.orc_pressure.restype = C.c_double
        L.orc_pres_lj_tail.restype = C.c_double
        L.orc_gaussian.restype = C.c_double
        L.orc_gaussian.argtypes = [C.c_void_p, C.c_double]
        L.orc_ran2.restype = C.c_double
        L.orc_ran2.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_gle_init.restype = C.c_int
        L.orc_rattle_s1.restype = C.c_int
        L.orc_rattle_s2.restype = C.c_int
        L.orc_evolve_rig_pi.restype = C.c_int

    # -- set-up --------------------------------------------------------------------------
    def setup_ewald(self, na: int, box, prm: RpmdParams) -> RpmdParams:
        b = _f64(box)
        self.lib.orc_setup_ewald(C.c_int(na), _d(b), C.byref(prm))
        return prm

    def setup_box_size(self, nm, rho, wmass, ncell):
        box = np.zeros(3)
        nc = (C.c_int * 3)(*ncell)
        self.lib.orc_setup_box_size(C.c_int(nm), C.c_double(rho), C.c_double(wmass), nc, _d(box))
        return box

    # -- forces ----------------------------------------------------------------------------
    def forces(self, prm: RpmdParams, r: np.ndarray, box, z, iopt: int):
        """forces(r,v,dvdr,nb,na,boxlxyz,z,virial,iopt); r has shape (nb, na, 3)."""
        r = _f64(r)
        nb, na = r.shape[0], r.shape[1]
        dvdr = np.zeros_like(r)
        vir = np.zeros(9)
        v = C.c_double()
        err = self.lib.orc_forces(C.byref(prm), _d(r), C.c_int(nb), C.c_int(na), _d(_f64(box)),
                                  _d(_f64(z)), C.c_int(iopt), C.byref(v), _d(dvdr), _d(vir))
        if err:
            raise RuntimeError(f"orc_forces error {err}")
        return v.value, dvdr, vir.reshape(3, 3).T.copy()

    def full_forces(self, prm: RpmdParams, r: np.ndarray, box, z):
        r = _f64(r)
        nb, na = r.shape[0], r.shape[1]
        dvdr = np.zeros_like(r)
        dvdr2 = np.zeros_like(r)
        vir = np.zeros(9)
        v, vew, voo, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        err = self.lib.orc_full_forces(C.byref(prm), _d(r), C.c_int(na), C.c_int(nb), C.byref(v),
                                       C.byref(vew), C.byref(voo), C.byref(vint), _d(vir),
                                       _d(_f64(z)), _d(_f64(box)), _d(dvdr), _d(dvdr2))
        if err:
            raise RuntimeError(f"orc_full_forces error {err}")
        return dict(v=v.value, vew=vew.value, voo=voo.value, vint=vint.value,
                    vir=vir.reshape(3, 3).T.copy(), dvdr=dvdr, dvdr2=dvdr2)

    def realft(self, data: np.ndarray, mode: int) -> np.ndarray:
        """data shape (n, m) = Fortran data(m, n); returns the transformed copy."""
        d = _f64(data).copy()
        n, m = d.shape
        err = self.lib.orc_realft(_d(d), C.c_int(m), C.c_int(n), C.c_int(mode))
        if err:
            raise RuntimeError(f"orc_realft error {err}")
        return d

    def ring_rpmd(self, mass3, nb, dt, beta):
        mon = np.zeros((nb, 4, 3))
        m = _f64(mass3)
        self.lib.orc_ring_rpmd(_d(m), _d(mon), C.c_int(3), C.c_int(nb), C.c_double(dt), C.c_double(beta))
        return mon

    def free_ring(self, p, r, monod):
        """the free ring-polymer update (normal-mode p, r of shape (nb, na, 3)); returns copies"""
        p, r = _f64(p).copy(), _f64(r).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_free_ring(_d(p), _d(r), _d(_f64(monod)), C.c_int(na), C.c_int(nb))
        return p, r

    # -- RNG / thermostats -------------------------------------------------------------------
    def rng(self, irun: int) -> RngState:
        s = RngState()
        self.lib.orc_rng_seed(C.byref(s), C.c_int(irun))
        return s

    def gaussians(self, s: RngState, n: int) -> np.ndarray:
        out = np.zeros(n)
        self.lib.orc_gaussian_fill(C.byref(s), _d(out), C.c_longlong(n))
        return out

    def sample_rpmd(self, nb, na, mass, beta, s: RngState) -> np.ndarray:
        p = np.zeros((nb, na, 3))
        self.lib.orc_sample_rpmd(_d(p), C.c_int(na), C.c_int(nb), _d(_f64(mass)), C.c_double(beta), C.byref(s))
        return p

    def therm_pro(self, p, mass, beta, ctau, dt, s: RngState):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_therm_pro(_d(p), _d(_f64(mass)), C.c_double(beta), C.c_double(ctau), C.c_double(dt),
                               C.c_int(na), C.c_int(nb), C.byref(s))
        return p

    def therm_pro_global(self, p, mass, beta, ctau, dt, s: RngState):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_therm_pro_global(_d(p), _d(_f64(mass)), C.c_double(beta), C.c_double(ctau), C.c_double(dt),
                                      C.c_int(na), C.c_int(nb), C.byref(s))
        return p

    def parinello_therm(self, p, mass, tau, dt, s: RngState, beta, nctot=0):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_parinello_therm(_d(p), _d(_f64(mass)), C.c_double(tau), C.c_int(na), C.c_int(nb),
                                     C.c_double(dt), C.byref(s), C.c_double(beta), C.c_int(nctot))
        return p

    def parinello_therm_loc(self, p, mass, tau, dt, s: RngState, beta):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_parinello_therm_loc(_d(p), _d(_f64(mass)), C.c_double(tau), C.c_int(na), C.c_int(nb),
                                         C.c_double(dt), C.byref(s), C.c_double(beta))
        return p

    def therm_gle(self, g: GleState, p, mass, s: RngState):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib.orc_therm_gle(C.byref(g), _d(p), _d(_f64(mass)), C.c_int(na), C.c_int(nb), C.byref(s))
        return p

    def gle_channels(self, g: GleState) -> np.ndarray:
        """the auxiliary momenta gp(3,na,nb,ns+1) as a (ns+1, nb, na, 3) array"""
        n = (g.ns + 1) * g.nb * g.na * 3
        return np.ctypeslib.as_array(g.gp, shape=(n,)).reshape(g.ns + 1, g.nb, g.na, 3).copy()

    # -- rigid molecules --------------------------------------------------------------------
    def rattle_s1(self, r, rold, p, mass, dt, dvdr, prm: RpmdParams):
        """one bead: r, rold, p, dvdr of shape (na, 3); returns (r, p, dvdr, vir, err)"""
        r, p, dvdr = _f64(r).copy(), _f64(p).copy(), _f64(dvdr).copy()
        vir = np.zeros(9)
        err = self.lib.orc_rattle_s1(_d(r), _d(_f64(rold)), _d(p), C.c_int(r.shape[0] // 3), _d(_f64(mass)),
                                     C.c_double(dt), _d(dvdr), _d(vir), C.c_double(prm.reoh), C.c_double(prm.theta))
        return r, p, dvdr, vir.reshape(3, 3).T.copy(), err

    def rattle_s2(self, r, p, mass, dt, dvdr, prm: RpmdParams):
        r, p, dvdr = _f64(r), _f64(p).copy(), _f64(dvdr).copy()
        vir = np.zeros(9)
        err = self.lib.orc_rattle_s2(_d(r), _d(p), C.c_int(r.shape[0] // 3), _d(_f64(mass)), C.c_double(dt),
                                     _d(dvdr), _d(vir), C.c_double(prm.reoh), C.c_double(prm.theta))
        return p, dvdr, vir.reshape(3, 3).T.copy(), err

    def bead_forces(self, r, beta, mass):
        r = _f64(r)
        nb, na = r.shape[0], r.shape[1]
        out = np.zeros_like(r)
        self.lib.orc_bead_forces(_d(r), C.c_int(na), C.c_int(nb), _d(out), C.c_double(beta), _d(_f64(mass)))
        return out

    def evolve_rig_pi(self, prm: RpmdParams, p, r, dvdr, dvdr2, dt, mass, box, z, beta, vir=None,
                      s: RngState | None = None):
        p, r = _f64(p).copy(), _f64(r).copy()
        dvdr, dvdr2 = _f64(dvdr).copy(), _f64(dvdr2).copy()
        nb, na = r.shape[0], r.shape[1]
        vir = np.zeros(9) if vir is None else _f64(np.asarray(vir).T).reshape(9).copy()
        vir_lf = np.zeros(9)
        v, vlf, vhf = C.c_double(), C.c_double(), C.c_double()
        if s is None:
            s = RngState()
            self.lib.orc_rng_init(C.byref(s))
        err = self.lib.orc_evolve_rig_pi(C.byref(prm), _d(p), _d(r), C.byref(v), C.byref(vlf), C.byref(vhf),
                                         _d(dvdr), _d(dvdr2), C.c_double(dt), _d(_f64(mass)), C.c_int(na),
                                         C.c_int(nb), _d(_f64(box)), _d(_f64(z)), C.c_double(beta), _d(vir),
                                         _d(vir_lf), C.byref(s))
        if err:
            raise RuntimeError(f"orc_evolve_rig_pi error {err}")
        return dict(p=p, r=r, dvdr=dvdr, dvdr2=dvdr2, v=v.value, v_lf=vlf.value, v_hf=vhf.value,
                    vir=vir.reshape(3, 3).T.copy(), vir_lf=vir_lf.reshape(3, 3).T.copy())

    def evolve_rig_pi_npt(self, prm: RpmdParams, p, r, dvdr, dvdr2, dt, mass, box, z, beta, vir, baro: "OrcBaro",
                          s: RngState | None = None):
        """evolve_rig_pi with nbaro = 1; vir (3x3) in/out as the reference's; returns the box too"""
        p, r = _f64(p).copy(), _f64(r).copy()
        dvdr, dvdr2 = _f64(dvdr).copy(), _f64(dvdr2).copy()
        nb, na = r.shape[0], r.shape[1]
        b = _f64(box).copy()
        w = _f64(np.asarray(vir).T).reshape(9).copy()
        vir_lf = np.zeros(9)
        v, vlf, vhf = C.c_double(), C.c_double(), C.c_double()
        if s is None:
            s = RngState()
            self.lib.orc_rng_init(C.byref(s))
        err = self.lib.orc_evolve_rig_pi_npt(C.byref(prm), _d(p), _d(r), C.byref(v), C.byref(vlf), C.byref(vhf),
                                             _d(dvdr), _d(dvdr2), C.c_double(dt), _d(_f64(mass)), C.c_int(na),
                                             C.c_int(nb), _d(b), _d(_f64(z)), C.c_double(beta), _d(w), _d(vir_lf),
                                             C.byref(s), C.byref(baro))
        if err:
            raise RuntimeError(f"orc_evolve_rig_pi_npt error {err}")
        return dict(p=p, r=r, dvdr=dvdr, dvdr2=dvdr2, v=v.value, v_lf=vlf.value, v_hf=vhf.value, box=b,
                    vir=w.reshape(3, 3).T.copy(), vir_lf=vir_lf.reshape(3, 3).T.copy())

    def matrix_exp(self, M, j=15, k=15):
        M = _f64(np.asarray(M).T)  # column-major
        n = M.shape[0]
        out = np.zeros(n * n)
        self.lib.orc_matrix_exp(_d(M), C.c_int(n), C.c_int(j), C.c_int(k), _d(out))
        return out.reshape(n, n).T.copy()

    def cholesky(self, A):
        A = _f64(np.asarray(A).T)
        n = A.shape[0]
        out = np.zeros(n * n)
        self.lib.orc_cholesky(_d(A), _d(out), C.c_int(n))
        return out.reshape(n, n).T.copy()

    def gle_init(self, A, ttau, na, nb, dt, s: RngState, beta) -> GleState:
        g = GleState()
        A = _f64(np.asarray(A).T)
        err = self.lib.orc_gle_init(C.byref(g), _d(A), C.c_int(A.shape[0] - 1), C.c_double(ttau),
                                    C.c_int(na), C.c_int(nb), C.c_double(dt), C.byref(s), C.c_double(beta))
        if err:
            raise RuntimeError(f"orc_gle_init error {err}")
        return g

    def gle_free(self, g: GleState):
        self.lib.orc_gle_free(C.byref(g))

    # -- the step ------------------------------------------------------------------------------
    def evolve_pi(self, prm: RpmdParams, p, r, dvdr, dvdr2, dt, mass, box, z, beta,
                  s: RngState | None = None, g: GleState | None = None):
        p, r = _f64(p).copy(), _f64(r).copy()
        dvdr, dvdr2 = _f64(dvdr).copy(), _f64(dvdr2).copy()
        nb, na = r.shape[0], r.shape[1]
        vir, vir_lf = np.zeros(9), np.zeros(9)
        v, vew, vlj, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        if s is None:
            s = RngState()
            self.lib.orc_rng_init(C.byref(s))
        err = self.lib.orc_evolve_pi(C.byref(prm), _d(p), _d(r), C.byref(v), C.byref(vew), C.byref(vlj),
                                     C.byref(vint), _d(dvdr), _d(dvdr2), C.c_double(dt), _d(_f64(mass)),
                                     C.c_int(na), C.c_int(nb), _d(_f64(box)), _d(_f64(z)),
                                     C.c_double(beta), _d(vir), _d(vir_lf), C.byref(s),
                                     C.byref(g) if g is not None else None)
        if err:
            raise RuntimeError(f"orc_evolve_pi error {err}")
        return dict(p=p, r=r, dvdr=dvdr, dvdr2=dvdr2, v=v.value, vew=vew.value, vlj=vlj.value,
                    vint=vint.value, vir=vir.reshape(3, 3).T.copy(), vir_lf=vir_lf.reshape(3, 3).T.copy())

    def evolve_pi_npt(self, prm: RpmdParams, p, r, dvdr, dvdr2, dt, mass, box, z, beta, vir_lf, baro: "OrcBaro",
                      s: RngState | None = None, g: GleState | None = None):
        """evolve_pi with nbaro = 1: baro is updated (MC counters); vir_lf (3x3) is the previous step's;
        returns the step's arrays, energies, box and virials"""
        p, r = _f64(p).copy(), _f64(r).copy()
        dvdr, dvdr2 = _f64(dvdr).copy(), _f64(dvdr2).copy()
        nb, na = r.shape[0], r.shape[1]
        b = _f64(box).copy()
        vir, vl = np.zeros(9), np.asarray(vir_lf, dtype=np.float64).T.reshape(9).copy()
        v, vew, vlj, vint = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        if s is None:
            s = RngState()
            self.lib.orc_rng_init(C.byref(s))
        err = self.lib.orc_evolve_pi_npt(C.byref(prm), _d(p), _d(r), C.byref(v), C.byref(vew), C.byref(vlj),
                                         C.byref(vint), _d(dvdr), _d(dvdr2), C.c_double(dt), _d(_f64(mass)),
                                         C.c_int(na), C.c_int(nb), _d(b), _d(_f64(z)), C.c_double(beta), _d(vir),
                                         _d(vl), C.byref(s), C.byref(g) if g is not None else None, C.byref(baro))
        if err:
            raise RuntimeError(f"orc_evolve_pi_npt error {err}")
        return dict(p=p, r=r, dvdr=dvdr, dvdr2=dvdr2, v=v.value, vew=vew.value, vlj=vlj.value, box=b,
                    vint=vint.value, vir=vir.reshape(3, 3).T.copy(), vir_lf=vl.reshape(3, 3).T.copy())

    # -- barostat helpers ----------------------------------------------------------------------
    def virial_ke(self, r, dvdr, dvdr2, beta, nctot=0):
        r = _f64(r)
        nb, na = r.shape[0], r.shape[1]
        tv, tq1, tq2 = C.c_double(), C.c_double(), C.c_double()
        tvxyz = np.zeros(3)
        self.lib.orc_virial_ke(_d(r), _d(_f64(dvdr)), _d(_f64(dvdr2)), C.byref(tv), _d(tvxyz), C.byref(tq1),
                               C.byref(tq2), C.c_double(beta), C.c_int(na), C.c_int(nb), C.c_int(nctot))
        return tv.value, tvxyz, tq1.value, tq2.value

    def pressure(self, prm: RpmdParams, vir, tv, na, box):
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        return self.lib.orc_pressure(C.byref(prm), _d(w), C.c_double(tv), C.c_int(na), _d(_f64(box)))

    def center_water(self, r):
        r = _f64(r)
        nb, nm = r.shape[0], r.shape[1] // 3
        pc = np.zeros((nm, 3))
        self.lib.orc_center_water(_d(r), _d(pc), C.c_int(nm), C.c_int(nb))
        return pc

    def beren_driver(self, prm: RpmdParams, baro: "OrcBaro", vir, tv, tvxyz, dt, r, box):
        r, b = _f64(r).copy(), _f64(box).copy()
        nb, na = r.shape[0], r.shape[1]
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        self.lib.orc_beren_driver(C.byref(prm), C.byref(baro), _d(w), C.c_double(tv), _d(_f64(tvxyz)),
                                  C.c_double(dt), _d(r), _d(b), C.c_int(na), C.c_int(nb))
        return r, b

    def mc_baro(self, prm: RpmdParams, baro: "OrcBaro", r, dvdr, dvdr2, vir, z, beta, box, s: RngState,
                eo: float = 0.0):
        r, g1, g2, b = _f64(r).copy(), _f64(dvdr).copy(), _f64(dvdr2).copy(), _f64(box).copy()
        nb, na = r.shape[0], r.shape[1]
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        eo = C.c_double(eo)
        err = self.lib.orc_mc_baro(C.byref(prm), C.byref(baro), _d(r), _d(g1), _d(g2), _d(w), C.byref(eo),
                                   _d(_f64(z)), C.c_double(beta), _d(b), C.c_int(na), C.c_int(nb), C.byref(s))
        if err:
            raise RuntimeError(f"orc_mc_baro error {err}")
        return dict(r=r, dvdr=g1, dvdr2=g2, vir=w.reshape(3, 3).T.copy(), eo=eo.value, box=b)

    def ran2(self, s: RngState, lo: float, hi: float) -> float:
        return self.lib.orc_ran2(C.byref(s), C.c_double(lo), C.c_double(hi))


class OrcBaro(C.Structure):
    """module barostat (src/module_barostat.f90:1-5): baro 1 BER, 2 MCI"""
    _fields_ = [("baro", C.c_int), ("taub", C.c_double), ("patm", C.c_double), ("dv", C.c_double),
                ("nacc", C.c_int), ("nmove", C.c_int)]


class RefPath:
    """The reference's own Fortran (force path, RNG, thermostat helpers) built from source."""

    def __init__(self, path: str = LIB_REF):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference is mounted")
        # RTLD_LAZY: realft_/rp_contract_nm_ (FFTW-dependent) stay unresolved and are never called
        self.lib = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        self.lib.gaussian_.restype = C.c_double
        self.lib.ran2_.restype = C.c_double

    def _common(self, name, ctype):
        return ctype.in_dll(self.lib, name)

    def set_params(self, prm: RpmdParams):
        """Fill the COMMON blocks the force path reads (src/main.f90:53-72)."""
        L = self.lib
        oo = (C.c_double * 4).in_dll(L, "oo_param_")
        oo[:] = [prm.oo_eps, prm.oo_sig, prm.oo_gam, prm.rcut]
        ew = (C.c_double * 4).in_dll(L, "ew_param_")
        ew[:] = [prm.alpha, prm.ecut, prm.wm, prm.wh]

        class EwSetup(C.Structure):
            _fields_ = [("wrcut", C.c_double), ("walpha", C.c_double), ("rkmax", C.c_double), ("kmax", C.c_int)]

        es = EwSetup.in_dll(L, "ew_setup_")
        es.wrcut, es.walpha, es.rkmax, es.kmax = prm.wrcut, prm.walpha, prm.rkmax, prm.kmax
        ip = (C.c_double * 4).in_dll(L, "intra_param_")
        ip[:] = [prm.apot, prm.bpot, prm.alp, prm.alpb]
        ge = (C.c_double * 2).in_dll(L, "geometry_")
        ge[:] = [prm.theta, prm.reoh]
        C.c_int.in_dll(L, "symmetry_").value = 1 if prm.iamcub else 0
        C.c_double.in_dll(L, "correct_").value = prm.sig
        (C.c_int * 3).in_dll(L, "rpmddft_")[:] = [0, 0, 0]
        C.c_int.in_dll(L, "epsr_").value = 0
        C.c_int.in_dll(L, "e3b_").value = 0

    def forces(self, prm: RpmdParams, r: np.ndarray, box, z, iopt: int):
        self.set_params(prm)
        r = _f64(r).copy()
        nb, na = r.shape[0], r.shape[1]
        dvdr = np.zeros_like(r)
        vir = np.zeros(9)
        v = C.c_double()
        b = _f64(box).copy()
        zz = _f64(z).copy()
        self.lib.forces_(_d(r), C.byref(v), _d(dvdr), C.byref(C.c_int(nb)), C.byref(C.c_int(na)), _d(b),
                         _d(zz), _d(vir), C.byref(C.c_int(iopt)))
        return v.value, dvdr, vir.reshape(3, 3).T.copy()

    def seed(self, irun: int):
        """src/main.f90:298-299: irun = -irun; test = gaussian(irun,1.d0).  gaussian keeps a
        cached second deviate across calls (src/random.f90:9); when one is pending the first call
        returns it without touching ran2 (idum stays negative), so draw again to re-seed."""
        self._idum = C.c_int(-irun)
        self.gaussian()
        if self._idum.value < 0:
            self.gaussian()

    def gaussian(self, sigma: float = 1.0) -> float:
        return self.lib.gaussian_(C.byref(self._idum), C.byref(C.c_double(sigma)))

    def set_nctot(self, nctot: int):
        """/constraint/ nctot, the constraint count parinello_therm subtracts from the degrees of
        freedom (therm_pro_global calls it too); 0 for flexible molecules"""
        (C.c_int * 2).in_dll(self.lib, "constraint_")[0] = nctot

    def sample_rpmd(self, nb, na, mass, beta):
        """thermostat::sample_rpmd (src/module_thermostat.f90:199-226) on the current ran2 stream"""
        p = np.zeros((nb, na, 3))
        self.lib._QMthermostatPsample_rpmd(_d(p), C.byref(C.c_int(na)), C.byref(C.c_int(nb)), _d(_f64(mass).copy()),
                                           C.byref(C.c_double(beta)), C.byref(self._idum))
        return p

    def pshift(self, p, mass):
        """thermostat::pshift (src/module_thermostat.f90:229-264)"""
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib._QMthermostatPpshift(_d(p), _d(_f64(mass).copy()), C.byref(C.c_int(na)), C.byref(C.c_int(nb)))
        return p

    def setup_ewald(self, na: int, box, prm: RpmdParams) -> dict:
        """setup_ewald (src/setup_options.f90:104-189): reads /ew_param/ and /symmetry/, writes
        /ew_setup/ and /symmetry/; returns wrcut, walpha, rkmax, kmax and iamcub"""
        self.set_params(prm)
        self.lib.setup_ewald_(C.byref(C.c_int(na)), _d(_f64(box).copy()))

        class EwSetup(C.Structure):
            _fields_ = [("wrcut", C.c_double), ("walpha", C.c_double), ("rkmax", C.c_double), ("kmax", C.c_int)]

        es = EwSetup.in_dll(self.lib, "ew_setup_")
        return dict(wrcut=es.wrcut, walpha=es.walpha, rkmax=es.rkmax, kmax=es.kmax,
                    iamcub=1 if C.c_int.in_dll(self.lib, "symmetry_").value else 0)

    def setup_box_size(self, lattice: str, rho: float, ncell, wmass: float):
        """setup_box_size (src/setup_options.f90:191-270) for the CUB/FCC/ICE lattices: returns
        (nm, boxlxyz); the cell counts go in through /lattice/"""
        lat = (C.c_int * 4).in_dll(self.lib, "lattice_")
        lat[0], lat[1], lat[2] = ncell
        C.c_int.in_dll(self.lib, "symmetry_").value = 1
        nm = C.c_int()
        box = np.zeros(3)
        name = C.create_string_buffer(lattice.encode()[:3].ljust(3), 3)
        # character(len=3) dummy: its length travels as a hidden trailing argument
        self.lib.setup_box_size_(name, C.byref(C.c_double(rho)), C.byref(nm), _d(box), C.byref(C.c_double(wmass)),
                                 C.c_size_t(3))
        return nm.value, box

    def therm_pro(self, p, mass, beta, ctau, dt):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib._QMglePtherm_pro(_d(p), _d(_f64(mass)), C.byref(C.c_double(beta)), C.byref(C.c_double(ctau)),
                                  C.byref(C.c_double(dt)), C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                  C.byref(self._idum))
        return p

    def parinello_therm(self, p, mass, tau, dt, beta, nctot=0):
        """thermostat::parinello_therm (src/module_thermostat.f90:6-54); nctot via /constraint/"""
        (C.c_int * 2).in_dll(self.lib, "constraint_")[0] = nctot
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib._QMthermostatPparinello_therm(_d(p), _d(_f64(mass)), C.byref(C.c_double(tau)), C.byref(C.c_int(na)),
                                               C.byref(C.c_int(nb)), C.byref(C.c_double(dt)), C.byref(self._idum),
                                               C.byref(C.c_double(beta)))
        return p

    def parinello_therm_loc(self, p, mass, tau, dt, beta):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib._QMthermostatPparinello_therm_loc(_d(p), _d(_f64(mass)), C.byref(C.c_double(tau)),
                                                   C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                                   C.byref(C.c_double(dt)), C.byref(self._idum),
                                                   C.byref(C.c_double(beta)))
        return p

    def therm_pro_global(self, p, mass, beta, ctau, dt):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        self.lib._QMglePtherm_pro_global(_d(p), _d(_f64(mass)), C.byref(C.c_double(beta)), C.byref(C.c_double(ctau)),
                                         C.byref(C.c_double(dt)), C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                         C.byref(self._idum))
        return p

    def gle_init(self, A, ttau, na, nb, dt, beta, workdir):
        """gle::therm_gle_init (src/module_gle.f90:146-227).  It reads GLE-A from the working
        directory and allocates module arrays, so it can run once per process."""
        A = np.asarray(A, dtype=np.float64)
        with open(os.path.join(workdir, "GLE-A"), "w") as f:
            f.write(f"{A.shape[0] - 1}\n")
            for row in A:
                f.write(" ".join(repr(float(x)) for x in row) + "\n")
        cwd = os.getcwd()
        os.chdir(workdir)
        try:
            self.lib._QMglePtherm_gle_init(C.byref(C.c_double(ttau)), C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                           C.byref(C.c_double(dt)), C.byref(self._idum), C.byref(C.c_double(beta)))
        finally:
            os.chdir(cwd)

    def therm_gle(self, p, mass):
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        dheat = C.c_double()
        self.lib._QMglePtherm_gle(_d(p), C.byref(dheat), _d(_f64(mass)), C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                  C.byref(self._idum))
        return p

    # -- rigid molecules (src/md_rattle.f90, src/evolve_rigid.f90) ---------------------------
    def rattle_setup(self, nm: int, prm: RpmdParams):
        """rattle_setup (src/md_rattle.f90:1-31) reads /geometry/, so the parameters go first"""
        self.set_params(prm)
        self.lib.rattle_setup_(C.byref(C.c_int(nm)))

    def rattle_s1(self, r, rold, p, mass, dt, dvdr):
        r, p, dvdr = _f64(r).copy(), _f64(p).copy(), _f64(dvdr).copy()
        na = r.shape[0]
        vir = np.zeros(9)
        self.lib.rattle_s1_(_d(r), _d(_f64(rold).copy()), _d(p), C.byref(C.c_int(na // 3)), _d(_f64(mass)),
                            C.byref(C.c_double(dt)), _d(dvdr), _d(vir), C.byref(C.c_int(na)))
        return r, p, dvdr, vir.reshape(3, 3).T.copy()

    def rattle_s2(self, r, p, mass, dt, dvdr):
        r, p, dvdr = _f64(r).copy(), _f64(p).copy(), _f64(dvdr).copy()
        na = r.shape[0]
        vir = np.zeros(9)
        self.lib.rattle_s2_(_d(r), _d(p), C.byref(C.c_int(na // 3)), _d(_f64(mass)), C.byref(C.c_double(dt)),
                            _d(dvdr), _d(vir), C.byref(C.c_int(na)))
        return p, dvdr, vir.reshape(3, 3).T.copy()

    def bead_forces(self, r, beta, mass):
        r = _f64(r).copy()
        nb, na = r.shape[0], r.shape[1]
        out = np.zeros_like(r)
        self.lib.bead_forces_(_d(r), C.byref(C.c_int(na)), C.byref(C.c_int(nb)), _d(out), C.byref(C.c_double(beta)),
                              _d(_f64(mass)))
        return out

    def matrix_exp(self, M, j=15, k=15):
        M = np.asfortranarray(M, dtype=np.float64)
        n = M.shape[0]
        out = np.zeros((n, n), order="F")
        self.lib._QMglePmatrix_exp(M.ctypes.data_as(_dp), C.byref(C.c_int(n)), C.byref(C.c_int(j)),
                                   C.byref(C.c_int(k)), out.ctypes.data_as(_dp))
        return np.array(out)

    def cholesky(self, A):
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        out = np.zeros((n, n), order="F")
        self.lib._QMglePcholesky(A.ctypes.data_as(_dp), out.ctypes.data_as(_dp), C.byref(C.c_int(n)))
        return np.array(out)

    def virial_ke(self, r, dvdr, dvdr2, beta, nctot=0):
        """virial_ke (src/kinetic.f90:1-54): (tv, tvxyz, tq1, tq2); nctot via /constraint/"""
        self.set_nctot(nctot)
        r, g1, g2 = _f64(r).copy(), _f64(dvdr).copy(), _f64(dvdr2).copy()
        nb, na = r.shape[0], r.shape[1]
        tv, tq1, tq2 = C.c_double(), C.c_double(), C.c_double()
        tvxyz = np.zeros(3)
        self.lib.virial_ke_(_d(r), _d(g1), _d(g2), C.byref(tv), _d(tvxyz), C.byref(tq1), C.byref(tq2),
                            C.byref(C.c_double(beta)), C.byref(C.c_int(na)), C.byref(C.c_int(nb)))
        return tv.value, tvxyz, tq1.value, tq2.value

    def set_barostat(self, baro: str, taub: float, patm: float, dv: float, nacc: int = 0, nmove: int = 0):
        """the variables of module barostat (src/module_barostat.f90:2-4)"""
        L = self.lib
        (C.c_char * 3).in_dll(L, "_QMbarostatEbaro").raw = baro.encode()[:3].ljust(3)
        C.c_double.in_dll(L, "_QMbarostatEtaub").value = taub
        C.c_double.in_dll(L, "_QMbarostatEpatm").value = patm
        C.c_double.in_dll(L, "_QMbarostatEdv").value = dv
        C.c_int.in_dll(L, "_QMbarostatEnacc").value = nacc
        C.c_int.in_dll(L, "_QMbarostatEnmove").value = nmove

    def barostat_counts(self):
        return (C.c_int.in_dll(self.lib, "_QMbarostatEnacc").value,
                C.c_int.in_dll(self.lib, "_QMbarostatEnmove").value)

    def pressure(self, prm: RpmdParams, vir, tv, na, box):
        """pressure (src/md_pressure.f90:1-35)"""
        self.set_params(prm)
        pres = C.c_double()
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        self.lib.pressure_(C.byref(pres), _d(w), C.byref(C.c_double(tv)), C.byref(C.c_int(na)), _d(_f64(box).copy()))
        return pres.value

    def center_water(self, r):
        """center_water (src/md_tools.f90:367-390)"""
        r = _f64(r).copy()
        nb, nm = r.shape[0], r.shape[1] // 3
        pc = np.zeros((nm, 3))
        self.lib.center_water_(_d(r), _d(pc), C.byref(C.c_int(nm)), C.byref(C.c_int(nb)))
        return pc

    def beren_driver(self, prm: RpmdParams, vir, tv, tvxyz, dt, r, box):
        """barostat::beren_driver (src/module_barostat.f90:8-26)"""
        self.set_params(prm)
        r, b = _f64(r).copy(), _f64(box).copy()
        nb, na = r.shape[0], r.shape[1]
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        self.lib._QMbarostatPberen_driver(_d(w), C.byref(C.c_double(tv)), _d(_f64(tvxyz).copy()),
                                          C.byref(C.c_double(dt)), _d(r), _d(b), C.byref(C.c_int(na)),
                                          C.byref(C.c_int(nb)))
        return r, b

    def mc_baro(self, prm: RpmdParams, r, dvdr, dvdr2, vir, z, beta, box, eo: float = 0.0):
        """barostat::mc_baro (src/module_barostat.f90:124-220) on the current ran2 stream (no RP
        contraction here: the reference's full_forces would call realft); eo in: the caller's v"""
        self.set_params(prm)
        r, g1, g2, b = _f64(r).copy(), _f64(dvdr).copy(), _f64(dvdr2).copy(), _f64(box).copy()
        nb, na = r.shape[0], r.shape[1]
        w = np.asarray(vir, dtype=np.float64).T.reshape(9).copy()
        eo = C.c_double(eo)
        (C.c_int * 2).in_dll(self.lib, "beaddiabatic_")[:] = [prm.nbdf1, prm.nbdf2]
        self.lib._QMbarostatPmc_baro(_d(r), _d(g1), _d(g2), _d(w), C.byref(eo), _d(_f64(z).copy()),
                                     C.byref(C.c_double(beta)), _d(b), C.byref(C.c_int(na)), C.byref(C.c_int(nb)),
                                     C.byref(self._idum))
        return dict(r=r, dvdr=g1, dvdr2=g2, vir=w.reshape(3, 3).T.copy(), eo=eo.value, box=b)

    def ran2(self, lo: float, hi: float) -> float:
        return self.lib.ran2_(C.byref(self._idum), C.byref(C.c_double(lo)), C.byref(C.c_double(hi)))

    def kinetic_rpmd(self, p, mass):
        """kinetic_rpmd (src/kinetic.f90:248-265)"""
        p = _f64(p).copy()
        nb, na = p.shape[0], p.shape[1]
        tk = C.c_double()
        self.lib.kinetic_rpmd_(_d(p), C.byref(C.c_int(na)), C.byref(tk), _d(_f64(mass).copy()), C.byref(C.c_int(nb)))
        return tk.value

    def ring_rpmd(self, mass3, nb, dt, beta):
        mon = np.zeros((nb, 4, 3))
        self.lib.ring_rpmd_(_d(_f64(mass3)), _d(mon), C.byref(C.c_int(3)), C.byref(C.c_int(nb)),
                            C.byref(C.c_double(dt)), C.byref(C.c_double(beta)))
        return mon
