"""ctypes wrapper of the FP64 oracle library (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the epic_amd package.

Arrays use the reference's Fortran layouts (see epic_oracle.h); numpy arrays
here are created in C order with the *reversed* shape so that the memory
layout is identical, e.g. a halo field f(-1:nz+1, -1:ny+1, -1:nx+1) is a
numpy array of shape (nx+3, ny+3, nz+3) indexed [ix+1, iy+1, iz+1].
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")


def build():
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        _declare(_lib)
    return _lib


class Dom(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lower", C.c_double * 3), ("extent", C.c_double * 3),
        ("glat", C.c_double), ("lambda_c", C.c_double), ("q0", C.c_double),
        ("f_cor", C.c_double * 3),
        ("zeta_zero", C.c_int * 2),
        ("vmin", C.c_double),
        ("nthreads", C.c_int),
        ("upper", C.c_double * 3), ("dx", C.c_double * 3), ("dxi", C.c_double * 3),
        ("vcell", C.c_double), ("center", C.c_double * 3), ("hl", C.c_double * 3),
        ("hli", C.c_double * 3),
    ]


class Parcels(C.Structure):
    _fields_ = [("n", C.c_int64)] + [
        (name, C.POINTER(C.c_double))
        for name in ("position", "vorticity", "B", "volume", "buoyancy", "humidity",
                     "delta_pos", "delta_vor", "delta_b", "strain", "Vetas", "Vtaus")
    ]


class Fields(C.Structure):
    _fields_ = [
        (name, C.POINTER(C.c_double))
        for name in ("velog", "vortg", "vtend", "velgradg", "tbuoyg", "dbuoyg", "humg", "volg")
    ] + [("nparg", C.POINTER(C.c_int)), ("nsparg", C.POINTER(C.c_int))]


class Time(C.Structure):
    _fields_ = [("fixed", C.c_int), ("fixed_dt", C.c_double), ("alpha", C.c_double), ("limit", C.c_double),
                ("precise_stop", C.c_int)]


class Timers(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("par2grid", "grid2par", "vor2vel", "vtend", "push")]


def _declare(L):
    P = C.POINTER
    d = C.c_double
    L.orc_dom_init.argtypes = [P(Dom)]
    L.orc_halo_size.argtypes = [P(Dom)]
    L.orc_halo_size.restype = C.c_int64
    L.orc_spec_size.argtypes = [P(Dom)]
    L.orc_spec_size.restype = C.c_int64
    L.orc_scherzinger_diagonalise.argtypes = [P(d), P(d), P(d)]
    L.orc_scherzinger_diagonalise.restype = None
    L.orc_scherzinger_eigenvalues.argtypes = [P(d), P(d)]
    L.orc_scherzinger_eigenvalues.restype = None
    L.orc_get_B33.argtypes = [P(d), d, P(d)]
    L.orc_get_points.argtypes = [P(d), P(d), d, P(d), P(d), P(d)]
    L.orc_get_dBdt.argtypes = [P(d), P(d), d, P(d)]
    L.orc_trilinear.argtypes = [P(Dom), P(d), P(C.c_int), P(d)]
    L.orc_trilinear.restype = None
    L.orc_get_index.argtypes = [P(Dom), P(d), P(C.c_int)]
    L.orc_get_index.restype = None
    L.orc_par2grid.argtypes = [P(Dom), P(Parcels), P(Fields), C.c_int]
    L.orc_grid2par.argtypes = [P(Dom), P(Parcels), P(Fields), C.c_int]
    for f in ("orc_forfft", "orc_revfft", "orc_dct", "orc_dst"):
        getattr(L, f).argtypes = [C.c_int, P(d)]
    L.orc_fftxyp2s.argtypes = [P(Dom), P(d), P(d)]
    L.orc_fftxys2p.argtypes = [P(Dom), P(d), P(d)]
    L.orc_diffx.argtypes = [P(Dom), P(d), P(d)]
    L.orc_diffx.restype = None
    L.orc_diffy.argtypes = [P(Dom), P(d), P(d)]
    L.orc_diffy.restype = None
    L.orc_inv_create.argtypes = [P(Dom), P(C.c_void_p)]
    L.orc_inv_destroy.argtypes = [C.c_void_p]
    L.orc_inv_destroy.restype = None
    L.orc_vor2vel.argtypes = [P(Dom), C.c_void_p, P(Fields)]
    L.orc_vel2vgrad.argtypes = [P(Dom), C.c_void_p, P(d), P(Fields)]
    L.orc_vorticity_tendency.argtypes = [P(Dom), C.c_void_p, P(Fields)]
    L.orc_apply_periodic_bc.argtypes = [P(Dom), P(d)]
    L.orc_apply_periodic_bc.restype = None
    L.orc_apply_reflective_bc.argtypes = [P(Dom), P(d), P(d)]
    L.orc_apply_reflective_bc.restype = None
    L.orc_ls_rk_step.argtypes = [P(Dom), C.c_void_p, P(Parcels), P(Fields), P(d), d, d, C.c_int]
    L.orc_get_time_step.argtypes = [P(Dom), P(Fields), d, P(Time), P(d), P(d), P(d)]
    L.orc_ls_rk_step_adaptive.argtypes = [P(Dom), C.c_void_p, P(Parcels), P(Fields), P(d), P(Time), d, C.c_int, P(d)]
    L.orc_parcel_diagnostics.argtypes = [P(Dom), P(Parcels), P(d), P(d)]
    L.orc_parcel_diagnostics.restype = None
    L.orc_field_diagnostics.argtypes = [P(Dom), P(Fields), P(d)]
    L.orc_field_diagnostics.restype = None
    L.orc_lsrk_linear.argtypes = [d, C.c_int]
    L.orc_lsrk_linear.restype = d
    L.orc_get_timers.argtypes = [P(Timers)]
    L.orc_get_timers.restype = None
    L.orc_reset_timers.argtypes = []
    L.orc_reset_timers.restype = None
    L.orc_default_amax.argtypes = [P(Dom)]
    L.orc_default_amax.restype = d
    L.orc_parcel_split.argtypes = [P(Dom), P(Parcels), C.c_int64, d, d, P(C.c_int64)]
    L.orc_parcel_merge.argtypes = [P(Dom), P(Parcels), P(C.c_int64)]
    L.orc_parcel_merge_stats.argtypes = [P(Dom), P(Parcels), P(C.c_int64), P(C.c_int64)]
    L.orc_apply_laplace.argtypes = [P(Dom), C.c_void_p, P(Parcels), P(Fields), C.c_int]
    L.orc_apply_gradient.argtypes = [P(Dom), P(Parcels), P(Fields), d, d, C.c_int]
    L.orc_volume_weighted_vorticity.argtypes = [P(Parcels), P(d)]
    L.orc_volume_weighted_vorticity.restype = None
    L.orc_apply_vortcor.argtypes = [P(Parcels), P(d)]
    L.orc_apply_vortcor.restype = None
    L.orc_nearest_ncell.argtypes = [P(Dom)]
    L.orc_nearest_ncell.restype = C.c_int64
    L.orc_nearest_cell_sort.argtypes = [P(Dom), C.c_int64, P(d), P(C.c_int), P(C.c_int), P(C.c_int),
                                        P(C.c_int)]


ERRORS = {
    -1: "par2grid: Wrong total number of parcels!",
    -2: "get_B33: Division by small number!",
    -3: "support point outside the halo grid",
    -4: "Only third and fourth order RK supported.",
    -5: "time step must be positive",
    -6: "FFT length not factorisable",
    -7: "allocation failed",
    -8: "Merge error: no near enough neighbour found.",
}


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}: {ERRORS.get(code, '?')}")
        self.code = code


def _chk(rc):
    if rc != 0:
        raise OracleError(rc)


def ptr(a, ctype=C.c_double):
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def make_dom(nx, ny, nz, lower, extent, glat=0.0, lambda_c=0.0, q0=0.0, f_cor=(0.0, 0.0, 0.0),
             zeta_zero=(False, False), vmin=0.0, nthreads=0):
    d = Dom()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.lower[:] = [float(v) for v in lower]
    d.extent[:] = [float(v) for v in extent]
    d.glat, d.lambda_c, d.q0 = glat, lambda_c, q0
    d.f_cor[:] = [float(v) for v in f_cor]
    d.zeta_zero[:] = [int(bool(v)) for v in zeta_zero]
    d.vmin = vmin
    d.nthreads = nthreads
    lib().orc_dom_init(C.byref(d))
    return d


# --------------------------------------------------------------------------
# small-array helpers
# --------------------------------------------------------------------------
def scherzinger_diagonalise(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    D = np.zeros(3)
    V = np.zeros(9)
    lib().orc_scherzinger_diagonalise(ptr(A), ptr(D), ptr(V))
    return D, V.reshape(3, 3).T.copy()   # V[:, j] = eigenvector j


def get_B33(B, vol):
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = C.c_double()
    _chk(lib().orc_get_B33(ptr(B), vol, C.byref(out)))
    return out.value


def get_points(pos, B, vol):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    ve, vt, pts = np.zeros(3), np.zeros(3), np.zeros(12)
    _chk(lib().orc_get_points(ptr(pos), ptr(B), vol, ptr(ve), ptr(vt), ptr(pts)))
    return ve, vt, pts.reshape(4, 3)


def get_dBdt(B, strain, vol):
    B = np.ascontiguousarray(B, dtype=np.float64)
    s = np.ascontiguousarray(strain, dtype=np.float64)
    out = np.zeros(5)
    _chk(lib().orc_get_dBdt(ptr(B), ptr(s), vol, ptr(out)))
    return out


def forfft(x):
    x = np.array(x, dtype=np.float64)
    _chk(lib().orc_forfft(len(x), ptr(x)))
    return x


def revfft(x):
    x = np.array(x, dtype=np.float64)
    _chk(lib().orc_revfft(len(x), ptr(x)))
    return x


def dct(x):
    """DCT-I of x[0..n] (length n+1)."""
    x = np.array(x, dtype=np.float64)
    _chk(lib().orc_dct(len(x) - 1, ptr(x)))
    return x


def dst(x):
    """DST-I of x[1..n] (length n); x[n] is returned as zero."""
    x = np.array(x, dtype=np.float64)
    _chk(lib().orc_dst(len(x), ptr(x)))
    return x


# --------------------------------------------------------------------------
# state: parcels + fields in the reference layouts
# --------------------------------------------------------------------------
class State:
    """Parcels (component-fastest) and halo fields of one member."""

    def __init__(self, dom: Dom, n: int):
        self.dom = dom
        self.n = int(n)
        nx, ny, nz = dom.nx, dom.ny, dom.nz
        self.hshape = (nx + 3, ny + 3, nz + 3)
        z = np.zeros
        self.position = z((n, 3))
        self.vorticity = z((n, 3))
        self.B = z((n, 5))
        self.volume = z(n)
        self.buoyancy = z(n)
        self.humidity = z(n)
        self.delta_pos = z((n, 3))
        self.delta_vor = z((n, 3))
        self.delta_b = z((n, 5))
        self.strain = z((n, 8))
        self.Vetas = z((n, 3))
        self.Vtaus = z((n, 3))
        hs = self.hshape
        self.velog = z((3,) + hs)
        self.vortg = z((3,) + hs)
        self.vtend = z((3,) + hs)
        self.velgradg = z((8,) + hs)
        self.tbuoyg = z(hs)
        self.dbuoyg = z(hs)
        self.humg = z(hs)
        self.volg = z(hs)
        self.nparg = np.zeros(hs, dtype=np.int32)
        self.nsparg = np.zeros(hs, dtype=np.int32)
        self._inv = None

    def _parcels(self):
        p = Parcels()
        p.n = self.n
        for name in ("position", "vorticity", "B", "volume", "buoyancy", "humidity", "delta_pos",
                     "delta_vor", "delta_b", "strain", "Vetas", "Vtaus"):
            setattr(p, name, ptr(getattr(self, name)))
        return p

    _PARCEL_ARRAYS = {"position": 3, "vorticity": 3, "B": 5, "volume": 0, "buoyancy": 0, "humidity": 0,
                      "delta_pos": 3, "delta_vor": 3, "delta_b": 5, "strain": 8, "Vetas": 3, "Vtaus": 3}

    def _grow(self, cap):
        for name, k in self._PARCEL_ARRAYS.items():
            a = getattr(self, name)
            b = np.zeros((cap, k) if k else cap)
            b[:self.n] = a[:self.n]
            setattr(self, name, b)

    def _trim(self):
        for name in self._PARCEL_ARRAYS:
            setattr(self, name, np.ascontiguousarray(getattr(self, name)[:self.n]))

    def parcel_split(self, lambda_max=4.0, amax=0.0):
        """parcel_split (parcel_split.f90:37-197); returns the number of splits."""
        cap = 2 * self.n
        self._grow(cap)
        ns = C.c_int64(0)
        p = self._parcels()
        rc = lib().orc_parcel_split(C.byref(self.dom), C.byref(p), cap, lambda_max, amax, C.byref(ns))
        self.n = int(p.n)
        self._trim()
        _chk(rc)
        return ns.value

    def parcel_merge(self):
        """parcel_merge (parcel_merge.f90, parcel_nearest.f90) on one rank;
        returns the number of merged small parcels.  self.merge_stats
        accumulates collect_merge_stats' counts (parcel_merge.f90:365-393):
        [n_big_close, n_way_parcel_mergers(1..7)]."""
        nm = C.c_int64(0)
        p = self._parcels()
        if not hasattr(self, "merge_stats"):
            self.merge_stats = np.zeros(8, dtype=np.int64)
        rc = lib().orc_parcel_merge_stats(C.byref(self.dom), C.byref(p), C.byref(nm),
                                          self.merge_stats.ctypes.data_as(C.POINTER(C.c_int64)))
        self.n = int(p.n)
        self._trim()
        _chk(rc)
        return nm.value

    def apply_laplace(self, l_reuse=None):
        """apply_laplace (parcel_correction.f90:142-181)."""
        flag = -1 if l_reuse is None else int(bool(l_reuse))
        _chk(lib().orc_apply_laplace(C.byref(self.dom), self.inv(), C.byref(self._parcels()),
                                     C.byref(self._fields()), flag))

    def apply_gradient(self, prefactor=1.8, max_compression=0.5, l_reuse=None):
        """apply_gradient (parcel_correction.f90:183-259)."""
        flag = -1 if l_reuse is None else int(bool(l_reuse))
        _chk(lib().orc_apply_gradient(C.byref(self.dom), C.byref(self._parcels()), C.byref(self._fields()),
                                      prefactor, max_compression, flag))

    def volume_weighted_vorticity(self):
        out = np.zeros(3)
        lib().orc_volume_weighted_vorticity(C.byref(self._parcels()), ptr(out))
        return out

    def apply_vortcor(self, vor_bar):
        vb = np.ascontiguousarray(vor_bar, dtype=np.float64)
        lib().orc_apply_vortcor(C.byref(self._parcels()), ptr(vb))

    def default_amax(self):
        return lib().orc_default_amax(C.byref(self.dom))

    def _fields(self):
        f = Fields()
        for name in ("velog", "vortg", "vtend", "velgradg", "tbuoyg", "dbuoyg", "humg", "volg"):
            setattr(f, name, ptr(getattr(self, name)))
        f.nparg = ptr(self.nparg, C.c_int)
        f.nsparg = ptr(self.nsparg, C.c_int)
        return f

    def inv(self):
        if self._inv is None:
            h = C.c_void_p()
            _chk(lib().orc_inv_create(C.byref(self.dom), C.byref(h)))
            self._inv = h
        return self._inv

    def __del__(self):
        if getattr(self, "_inv", None) is not None and _lib is not None:
            _lib.orc_inv_destroy(self._inv)
            self._inv = None

    # reference call surface
    def par2grid(self, l_reuse=None):
        flag = -1 if l_reuse is None else int(bool(l_reuse))
        _chk(lib().orc_par2grid(C.byref(self.dom), C.byref(self._parcels()), C.byref(self._fields()), flag))

    def grid2par(self, add=None):
        flag = -1 if add is None else int(bool(add))
        _chk(lib().orc_grid2par(C.byref(self.dom), C.byref(self._parcels()), C.byref(self._fields()), flag))

    def vor2vel(self):
        _chk(lib().orc_vor2vel(C.byref(self.dom), self.inv(), C.byref(self._fields())))

    def vel2vgrad(self, svel):
        svel = np.ascontiguousarray(svel, dtype=np.float64)
        _chk(lib().orc_vel2vgrad(C.byref(self.dom), self.inv(), ptr(svel), C.byref(self._fields())))

    def vorticity_tendency(self):
        _chk(lib().orc_vorticity_tendency(C.byref(self.dom), self.inv(), C.byref(self._fields())))

    def get_time_step(self, t, alpha=0.2, limit=0.0, precise_stop=False, fixed_dt=None):
        """get_time_step (rk_utils.f90:73-183) from the current velgradg and
        tbuoyg; returns (dt, gmax, bmax)."""
        tm = Time(1 if fixed_dt is not None else 0, fixed_dt or 0.0, alpha, limit, int(bool(precise_stop)))
        dt, g, b = C.c_double(0), C.c_double(0), C.c_double(0)
        _chk(lib().orc_get_time_step(C.byref(self.dom), C.byref(self._fields()), t, C.byref(tm), C.byref(dt),
                                     C.byref(g), C.byref(b)))
        return dt.value, g.value, b.value

    def ls_rk_step_adaptive(self, t, alpha=0.2, limit=0.0, precise_stop=False, t_initial=0.0, rk_order=4):
        """ls_rk_step with get_time_step (ls_rk.f90:88-136); returns (t, dt)."""
        tm = Time(0, 0.0, alpha, limit, int(bool(precise_stop)))
        tt, dt = C.c_double(t), C.c_double(0)
        _chk(lib().orc_ls_rk_step_adaptive(C.byref(self.dom), self.inv(), C.byref(self._parcels()),
                                           C.byref(self._fields()), C.byref(tt), C.byref(tm), t_initial, rk_order,
                                           C.byref(dt)))
        return tt.value, dt.value

    PARCEL_STATS = ("ape", "ke", "n_small", "avg_lam", "avg_vol", "std_lam", "std_vol", "sum_vol", "rms_xi",
                    "rms_eta", "rms_zeta", "ntot_par", "enstrophy", "nsplits", "nbig_iclo", "nmerges",
                    "min_buoy", "max_buoy")
    FIELD_STATS = ("rms_v", "avg_npar", "avg_nspar", "keg", "apeg", "eng", "abserr_v", "max_npar", "max_buoy",
                   "min_npar", "min_buoy")

    def parcel_diagnostics(self, counters=(0, 0, 0)):
        """calculate_parcel_diagnostics (parcel_diagnostics.f90:51-159); the
        velocity is delta_pos (after grid2par)."""
        out = np.zeros(18)
        c = np.asarray(counters, dtype=np.float64)
        lib().orc_parcel_diagnostics(C.byref(self.dom), C.byref(self._parcels()), ptr(c), ptr(out))
        return dict(zip(self.PARCEL_STATS, out))

    def field_diagnostics(self):
        """calculate_field_diagnostics (field_diagnostics.f90:40-154)."""
        out = np.zeros(11)
        lib().orc_field_diagnostics(C.byref(self.dom), C.byref(self._fields()), ptr(out))
        return dict(zip(self.FIELD_STATS, out))

    def ls_rk_step(self, t, dt, t_initial=0.0, rk_order=4):
        tt = C.c_double(t)
        _chk(lib().orc_ls_rk_step(C.byref(self.dom), self.inv(), C.byref(self._parcels()),
                                  C.byref(self._fields()), C.byref(tt), dt, t_initial, rk_order))
        return tt.value

    def fftxyp2s(self, fp):
        d = self.dom
        fs = np.zeros((d.nx, d.ny, d.nz + 1))
        fp = np.ascontiguousarray(fp, dtype=np.float64)
        _chk(lib().orc_fftxyp2s(C.byref(d), ptr(fp), ptr(fs)))
        return fs

    def fftxys2p(self, fs, fp=None):
        d = self.dom
        if fp is None:
            fp = np.zeros(self.hshape)
        fs = np.ascontiguousarray(fs, dtype=np.float64)
        _chk(lib().orc_fftxys2p(C.byref(d), ptr(fs), ptr(fp)))
        return fp

    def diffx(self, fs):
        ds = np.zeros_like(fs)
        lib().orc_diffx(C.byref(self.dom), ptr(np.ascontiguousarray(fs)), ptr(ds))
        return ds

    def diffy(self, fs):
        ds = np.zeros_like(fs)
        lib().orc_diffy(C.byref(self.dom), ptr(np.ascontiguousarray(fs)), ptr(ds))
        return ds


def nearest_cell_sort(dom, position):
    """Reference counting sort: returns (loca, nppc, kc1, node) as in
    parcel_nearest.f90 (loca/kc1 1-based, node 0-based parcel indices)."""
    position = np.ascontiguousarray(position, dtype=np.float64)
    n = position.shape[0]
    ncell = lib().orc_nearest_ncell(C.byref(dom))
    loca = np.zeros(n, dtype=np.int32)
    nppc = np.zeros(ncell, dtype=np.int32)
    kc1 = np.zeros(ncell, dtype=np.int32)
    node = np.zeros(n, dtype=np.int32)
    _chk(lib().orc_nearest_cell_sort(C.byref(dom), n, ptr(position), ptr(loca, C.c_int), ptr(nppc, C.c_int),
                                     ptr(kc1, C.c_int), ptr(node, C.c_int)))
    return loca, nppc, kc1, node


def timers():
    t = Timers()
    lib().orc_get_timers(C.byref(t))
    return {n: getattr(t, n) for n, _ in Timers._fields_}


def reset_timers():
    lib().orc_reset_timers()
