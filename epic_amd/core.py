"""Host-side mirror of EPIC's call surface over the C-ABI of libepic_gpu.so.

The method names follow the reference procedures they replace:
``par2grid(l_reuse)`` (parcel_interpl.f90:146), ``grid2par(add)`` (:405),
``vor2vel()`` (inversion.f90:25), ``vorticity_tendency()`` (:321) and
``ls_rk_step(t)`` (ls_rk.f90:88).  Arrays are numpy float64 in the same
layouts as the oracle's State: parcels (n, 3)/(n, 5)/(n,), fields
(ncomp, nx+3, ny+3, nz+3) indexed [c, ix+1, iy+1, iz+1].

The compute path is the HIP library only: importing this module fails loudly
when libepic_gpu.so is missing.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import re
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libepic_gpu.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "epic_gpu.h")

FIELDS = {"volg": (0, 1), "vortg": (1, 3), "tbuoyg": (2, 1), "dbuoyg": (3, 1), "humg": (4, 1),
          "velog": (5, 3), "velgradg": (6, 8), "vtend": (7, 3), "nparg": (8, 1), "nsparg": (9, 1)}

ERRORS = {-1: "EPIC_E_INVALID", -2: "EPIC_E_HIP", -3: "EPIC_E_PARCEL_COUNT", -4: "EPIC_E_B33",
          -5: "EPIC_E_OUT_OF_GRID", -6: "EPIC_E_RK_ORDER", -7: "EPIC_E_DT", -8: "EPIC_E_NOMEM",
          -9: "EPIC_E_CAPACITY", -10: "EPIC_E_MERGE"}


class EpicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lower", C.c_double * 3),
                ("extent", C.c_double * 3)]


class Phys(C.Structure):
    _fields_ = [("glat", C.c_double), ("lambda_c", C.c_double), ("q0", C.c_double), ("f_cor", C.c_double * 3)]


class Flags(C.Structure):
    _fields_ = [("zeta_bndry_zero", C.c_int * 2), ("rk_order", C.c_int), ("vmin", C.c_double)]


class TimeOpts(C.Structure):
    """epic_time_t (options.f90:99-107)."""
    _fields_ = [("fixed", C.c_int), ("fixed_dt", C.c_double), ("alpha", C.c_double), ("limit", C.c_double),
                ("precise_stop", C.c_int)]


_lib = None

# handles still open when the interpreter exits are destroyed by an atexit
# hook, while the HIP runtime (and a profiler's tool library around it) is
# still up, never by garbage collection during interpreter teardown
_live_batches = weakref.WeakSet()
_live_members = weakref.WeakSet()


@atexit.register
def _close_live_handles():
    for b in list(_live_batches):
        b.close()
    for m in list(_live_members):
        m.close()


def load_library(path=None):
    """Load libepic_gpu.so (or the build named by EPIC_GPU_LIB); raises if it
    has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("EPIC_GPU_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m epic_amd.build`")
    L = C.CDLL(path)
    P = C.POINTER
    d, i, i64 = C.c_double, C.c_int, C.c_int64
    dp, ip = P(d), P(i)
    h = C.c_void_p
    sig = {
        "epic_create": ([P(Grid), P(Phys), P(Flags), i64, P(h)], i),
        "epic_destroy": ([h], i),
        "epic_last_error": ([h], C.c_char_p),
        "epic_set_parcels": ([h, i64, dp, dp, dp, dp, dp, dp], i),
        "epic_get_parcels": ([h, i64, dp, dp, dp, dp, dp, dp], i),
        "epic_get_parcel_tendencies": ([h, i64, dp, dp, dp], i),
        "epic_set_parcel_tendencies": ([h, i64, dp, dp], i),
        "epic_get_field": ([h, i, dp], i),
        "epic_set_field": ([h, i, dp], i),
        "epic_get_field_int": ([h, i, ip], i),
        "epic_par2grid": ([h, i], i),
        "epic_vor2vel": ([h], i),
        "epic_vorticity_tendency": ([h], i),
        "epic_grid2par": ([h, i], i),
        "epic_ls_rk_step": ([h, dp, d, d], i),
        "epic_get_time_step": ([h, d, P(TimeOpts), dp, dp, dp], i),
        "epic_ls_rk_step_adaptive": ([h, dp, P(TimeOpts), d, dp], i),
        "epic_parcel_diagnostics": ([h, dp, i], i),
        "epic_parcel_diagnostics_noalloc": ([h, dp, i], i),
        "epic_resort": ([h], i),
        "epic_get_config": ([h, P(Grid), P(Phys), P(Flags)], i),
        "epic_write_parcels_nc": ([h, C.c_char_p, d], i),
        "epic_read_parcels_nc": ([h, C.c_char_p, dp], i),
        "epic_write_fields_nc": ([h, C.c_char_p, d], i),
        "epic_nc_write_parcels": ([C.c_char_p, P(Grid), P(Phys), P(Flags), i64, dp, dp, dp, dp, dp, dp, d], i),
        "epic_nc_read_parcels": ([C.c_char_p, i64, P(i64), P(Grid), dp, dp, dp, dp, dp, dp, dp], i),
        "epic_nc_write_fields": ([C.c_char_p, P(Grid), P(Phys), P(Flags), d] + [dp] * 9, i),
        "epic_nc_last_error": ([], C.c_char_p),
        "epic_set_resort_interval": ([h, i], i),
        "epic_field_diagnostics": ([h, dp], i),
        "epic_cell_sort": ([h, ip, ip, ip, ip], i),
        "epic_nearest_ncell": ([h], i64),
        "epic_num_parcels": ([h], i64),
        "epic_set_profiling": ([h, i], i),
        "epic_get_timers": ([h, dp, ip, i], i),
        "epic_member_bytes": ([P(Grid), i64], i64),
        "epic_parcel_split": ([h, d, d, P(i64)], i),
        "epic_parcel_merge": ([h, P(i64)], i),
        "epic_parcel_merge_stats": ([h, P(i64)], i),
        "epic_apply_laplace": ([h, i], i),
        "epic_apply_gradient": ([h, d, d, i], i),
        "epic_init_parcel_correction": ([h, dp], i),
        "epic_apply_vortcor": ([h], i),
        "epic_autotune": ([h, ip, ip], i),
        "epic_set_tiling": ([h, i, i], i),
        "epic_batch_create": ([P(h), i, P(h)], i),
        "epic_batch_destroy": ([h], i),
        "epic_batch_last_error": ([h], C.c_char_p),
        "epic_batch_step": ([h, dp, dp, dp, ip], i),
        "epic_batch_prepare": ([h], i),
        "epic_batch_launch": ([h, dp, dp, dp], i),
        "epic_batch_wait": ([h, dp, ip], i),
        "epic_nc_file_info": ([C.c_char_p, C.c_char_p, dp, P(i64), P(Grid)], i),
        "epic_nc_read_fields": ([C.c_char_p, i, P(Grid), dp, dp, dp, dp], i),
        "epic_read_fields_nc": ([h, C.c_char_p, i, dp], i),
        "epic_parcel_default": ([h, i, d], i),
        "epic_init_parcels_from_grids": ([h], i),
        "epic_next_write_time": ([d, d], d),
        "epic_setup_restart": ([h, C.c_char_p, dp, C.c_char_p], i),
        "epic_diagonalise": ([i64, dp, dp, dp], i),
        "epic_diagonalise_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def header_symbols(path=HEADER):
    """Names of the functions declared in include/epic_gpu.h."""
    text = open(path).read()
    return sorted(set(re.findall(r"\b(epic_[a-z0-9_]+)\s*\(", text)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _opt(flag):
    return -1 if flag is None else int(bool(flag))


class Member:
    """One EPIC domain resident on the GPU."""

    def __init__(self, nx, ny, nz, lower, extent, glat=0.0, lambda_c=0.0, q0=0.0, f_cor=(0.0, 0.0, 0.0),
                 zeta_zero=(False, False), rk_order=4, vmin=0.0, max_parcels=0):
        L = load_library()
        g = Grid(nx, ny, nz, (C.c_double * 3)(*lower), (C.c_double * 3)(*extent))
        p = Phys(glat, lambda_c, q0, (C.c_double * 3)(*f_cor))
        f = Flags((C.c_int * 2)(*[int(bool(z)) for z in zeta_zero]), rk_order, vmin)
        self._h = C.c_void_p()
        rc = L.epic_create(C.byref(g), C.byref(p), C.byref(f), int(max_parcels), C.byref(self._h))
        if rc != 0:
            raise EpicError(rc, "epic_create failed")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = 0
        _live_members.add(self)

    @classmethod
    def from_workload(cls, mem, rk_order=4, max_parcels=None):
        m = cls(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, mem.glat, mem.lambda_c, mem.q0, mem.f_cor,
                mem.zeta_zero, rk_order=rk_order, max_parcels=max_parcels or mem.n)
        m.set_parcels(mem.position, mem.vorticity, mem.B, mem.volume, mem.buoyancy, mem.humidity)
        return m

    def _chk(self, rc):
        if rc != 0:
            raise EpicError(rc, self.last_error())

    def last_error(self):
        msg = load_library().epic_last_error(self._h)
        return msg.decode() if msg else ""

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().epic_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def set_parcels(self, position, vorticity, B, volume, buoyancy, humidity):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (position, vorticity, B, volume, buoyancy, humidity)]
        n = arrs[0].shape[0]
        self._chk(load_library().epic_set_parcels(self._h, n, *[_dp(a) for a in arrs]))
        self.n = n

    def get_parcels(self):
        n = self.n
        out = {"position": np.zeros((n, 3)), "vorticity": np.zeros((n, 3)), "B": np.zeros((n, 5)),
               "volume": np.zeros(n), "buoyancy": np.zeros(n), "humidity": np.zeros(n)}
        self._chk(load_library().epic_get_parcels(self._h, n, *[_dp(out[k]) for k in
                                                               ("position", "vorticity", "B", "volume", "buoyancy",
                                                                "humidity")]))
        return out

    def get_tendencies(self):
        n = self.n
        dp, dv, st = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 8))
        self._chk(load_library().epic_get_parcel_tendencies(self._h, n, _dp(dp), _dp(dv), _dp(st)))
        return dp, dv, st

    def set_tendencies(self, delta_pos, delta_vor):
        a = np.ascontiguousarray(delta_pos, dtype=np.float64)
        b = np.ascontiguousarray(delta_vor, dtype=np.float64)
        self._chk(load_library().epic_set_parcel_tendencies(self._h, self.n, _dp(a), _dp(b)))

    def get_field(self, name):
        which, nc = FIELDS[name]
        shape = (nc, self.nx + 3, self.ny + 3, self.nz + 3)
        if name in ("nparg", "nsparg"):
            out = np.zeros(shape[1:], dtype=np.int32)
            self._chk(load_library().epic_get_field_int(self._h, which, out.ctypes.data_as(C.POINTER(C.c_int))))
            return out
        out = np.zeros(shape)
        self._chk(load_library().epic_get_field(self._h, which, _dp(out)))
        return out[0] if nc == 1 else out

    def set_field(self, name, values):
        which, nc = FIELDS[name]
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(nc, self.nx + 3, self.ny + 3, self.nz + 3)
        self._chk(load_library().epic_set_field(self._h, which, _dp(a)))

    # ---- reference call surface ----
    def par2grid(self, l_reuse=None):
        self._chk(load_library().epic_par2grid(self._h, _opt(l_reuse)))

    def vor2vel(self):
        self._chk(load_library().epic_vor2vel(self._h))

    def vorticity_tendency(self):
        self._chk(load_library().epic_vorticity_tendency(self._h))

    def grid2par(self, add=None):
        self._chk(load_library().epic_grid2par(self._h, _opt(add)))

    def ls_rk_step(self, t, dt, t_initial=0.0):
        tt = C.c_double(t)
        self._chk(load_library().epic_ls_rk_step(self._h, C.byref(tt), dt, t_initial))
        return tt.value

    def get_time_step(self, t, alpha=0.2, limit=0.0, precise_stop=False, fixed_dt=None):
        """get_time_step (rk_utils.f90:73-183) from the current fields;
        returns (dt, gmax, bmax)."""
        o = TimeOpts(1 if fixed_dt is not None else 0, fixed_dt or 0.0, alpha, limit, int(bool(precise_stop)))
        dt, g, b = C.c_double(0), C.c_double(0), C.c_double(0)
        self._chk(load_library().epic_get_time_step(self._h, t, C.byref(o), C.byref(dt), C.byref(g), C.byref(b)))
        return dt.value, g.value, b.value

    def ls_rk_step_adaptive(self, t, alpha=0.2, limit=0.0, precise_stop=False, t_initial=0.0):
        """ls_rk_step with get_time_step (ls_rk.f90:88-136); returns (t, dt)."""
        o = TimeOpts(0, 0.0, alpha, limit, int(bool(precise_stop)))
        tt, dt = C.c_double(t), C.c_double(0)
        self._chk(load_library().epic_ls_rk_step_adaptive(self._h, C.byref(tt), C.byref(o), t_initial, C.byref(dt)))
        return tt.value, dt.value

    PARCEL_STATS = ("ape", "ke", "n_small", "avg_lam", "avg_vol", "std_lam", "std_vol", "sum_vol", "rms_xi",
                    "rms_eta", "rms_zeta", "ntot_par", "enstrophy", "nsplits", "nbig_iclo", "nmerges",
                    "min_buoy", "max_buoy")
    FIELD_STATS = ("rms_v", "avg_npar", "avg_nspar", "keg", "apeg", "eng", "abserr_v", "max_npar", "max_buoy",
                   "min_npar", "min_buoy")

    def parcel_diagnostics(self, reset_counters=False, noalloc=False):
        """calculate_parcel_diagnostics (parcel_diagnostics.f90:51-159);
        noalloc: through epic_parcel_diagnostics_noalloc (no device or pinned
        allocation, a fixed per-GPU scratch array)."""
        out = np.zeros(18)
        L = load_library()
        fn = L.epic_parcel_diagnostics_noalloc if noalloc else L.epic_parcel_diagnostics
        self._chk(fn(self._h, _dp(out), int(bool(reset_counters))))
        return dict(zip(self.PARCEL_STATS, out))

    def field_diagnostics(self):
        """calculate_field_diagnostics (field_diagnostics.f90:40-154)."""
        out = np.zeros(11)
        self._chk(load_library().epic_field_diagnostics(self._h, _dp(out)))
        return dict(zip(self.FIELD_STATS, out))

    TIMERS = ("par2grid", "cell_sort", "vor2vel", "vorticity_tendency", "push", "grid2par", "axes", "p2g_tiles")

    def set_profiling(self, on=True, graph=False):
        """Per-phase GPU timers: eager launches between event pairs, or (graph)
        the captured step graph replayed with device wall-clock stamps
        around every phase."""
        self._chk(load_library().epic_set_profiling(self._h, (2 if graph else 1) if on else 0))

    def timers(self, reset=True):
        ms = np.zeros(len(self.TIMERS))
        calls = np.zeros(len(self.TIMERS), dtype=np.int32)
        self._chk(load_library().epic_get_timers(self._h, _dp(ms), calls.ctypes.data_as(C.POINTER(C.c_int)),
                                                 int(bool(reset))))
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(self.TIMERS)}

    def cell_sort(self):
        L = load_library()
        ncell = L.epic_nearest_ncell(self._h)
        loca = np.zeros(self.n, dtype=np.int32)
        node = np.zeros(self.n, dtype=np.int32)
        nppc = np.zeros(ncell, dtype=np.int32)
        kc1 = np.zeros(ncell, dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))   # noqa: E731
        self._chk(L.epic_cell_sort(self._h, ip(loca), ip(nppc), ip(kc1), ip(node)))
        return loca, nppc, kc1, node

    # ---- parcel operations ----
    def parcel_split(self, lambda_max=4.0, amax=0.0):
        """parcel_split (parcel_split.f90:37-197); returns the number of splits."""
        ns = C.c_int64(0)
        self._chk(load_library().epic_parcel_split(self._h, lambda_max, amax, C.byref(ns)))
        self.n = int(load_library().epic_num_parcels(self._h))
        return ns.value

    def parcel_merge_stats(self):
        """n_way_parcel_mergers (parcel_merge.f90:381-391) since the last
        diagnostics reset: entry j-1 counts the targets that took j small
        parcels (the last entry: seven or more)."""
        out = np.zeros(7, dtype=np.int64)
        self._chk(load_library().epic_parcel_merge_stats(self._h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def parcel_merge(self):
        """parcel_merge (parcel_merge.f90) on one rank; returns the number of
        parcels merged away."""
        nm = C.c_int64(0)
        self._chk(load_library().epic_parcel_merge(self._h, C.byref(nm)))
        self.n = int(load_library().epic_num_parcels(self._h))
        return nm.value

    # ---- parcel correction (parcel_correction.f90) ----
    def apply_laplace(self, l_reuse=None):
        """apply_laplace (parcel_correction.f90:142-181)."""
        self._chk(load_library().epic_apply_laplace(self._h, _opt(l_reuse)))

    def apply_gradient(self, prefactor=1.8, max_compression=0.5, l_reuse=None):
        """apply_gradient (parcel_correction.f90:183-259); defaults of
        options.f90:92-93."""
        self._chk(load_library().epic_apply_gradient(self._h, prefactor, max_compression, _opt(l_reuse)))

    def init_parcel_correction(self):
        """init_parcel_correction (parcel_correction.f90:42-83); returns the
        stored volume-weighted mean vorticity."""
        out = np.zeros(3)
        self._chk(load_library().epic_init_parcel_correction(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def apply_vortcor(self):
        """apply_vortcor (parcel_correction.f90:89-138)."""
        self._chk(load_library().epic_apply_vortcor(self._h))

    # ---- NetCDF files (NetCDF classic format, the reference's schema) ----
    def write_parcels_nc(self, path, t):
        """write_netcdf_parcels (parcel_netcdf.f90:201-309)."""
        self._chk(load_library().epic_write_parcels_nc(self._h, str(path).encode(), t))

    def read_parcels_nc(self, path):
        """read_netcdf_parcels (parcel_netcdf.f90:311-724); returns the time."""
        t = C.c_double(0)
        self._chk(load_library().epic_read_parcels_nc(self._h, str(path).encode(), C.byref(t)))
        self.n = int(load_library().epic_num_parcels(self._h))
        return t.value

    def write_fields_nc(self, path, t):
        """write_netcdf_fields (field_netcdf.f90:182-266)."""
        self._chk(load_library().epic_write_fields_nc(self._h, str(path).encode(), t))

    def read_fields_nc(self, path, step=-1):
        """read_netcdf_fields (field_netcdf.f90:306-386) into vortg, tbuoyg,
        humg; returns the record's time."""
        t = C.c_double(0)
        self._chk(load_library().epic_read_fields_nc(self._h, str(path).encode(), int(step), C.byref(t)))
        return t.value

    # ---- parcel set-up and restart (parcel_init.f90, utils.f90) ----
    def parcel_default(self, n_per_cell=8, lambda_max=4.0):
        """parcel_default (parcel_init.f90:31-100)."""
        self._chk(load_library().epic_parcel_default(self._h, int(n_per_cell), float(lambda_max)))
        self.n = int(load_library().epic_num_parcels(self._h))

    def init_parcels_from_grids(self):
        """init_parcels_from_grids (parcel_init.f90:174-203)."""
        self._chk(load_library().epic_init_parcels_from_grids(self._h))

    def setup_restart(self, path):
        """setup_restart + the restart branch of setup_fields_and_parcels
        (utils.f90:144-216); returns (t, file_type)."""
        t = C.c_double(0)
        ft = C.create_string_buffer(16)
        self._chk(load_library().epic_setup_restart(self._h, str(path).encode(), C.byref(t), ft))
        self.n = int(load_library().epic_num_parcels(self._h))
        return t.value, ft.value.decode()

    # ---- physical re-sort ----
    def resort(self):
        """Reorder the parcel arrays into cell order (not observable)."""
        self._chk(load_library().epic_resort(self._h))

    def set_resort_interval(self, steps):
        self._chk(load_library().epic_set_resort_interval(self._h, int(steps)))

    # ---- tiling ----
    def autotune(self):
        """Times the tile z chunks on this member; returns (tz_p2g, tz_g2p)."""
        a, b = C.c_int(0), C.c_int(0)
        self._chk(load_library().epic_autotune(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tiling(self, tz_p2g, tz_g2p):
        self._chk(load_library().epic_set_tiling(self._h, int(tz_p2g), int(tz_g2p)))


def member_bytes(nx, ny, nz, max_parcels):
    """Device bytes a member of this grid and capacity holds."""
    g = Grid(nx, ny, nz, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(1, 1, 1))
    return int(load_library().epic_member_bytes(C.byref(g), int(max_parcels)))


def diagonalise(A):
    """scherzinger_diagonalise on the device for symmetric matrices A (n, 3,
    3); returns D (n, 3) descending and V (n, 3, 3) with V[i, :, k] the
    eigenvector of D[i, k]."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    # Fortran A(3,3,n): element (r, c, i) at 9 i + r + 3 c
    a = np.ascontiguousarray(np.transpose(A, (0, 2, 1))).reshape(-1)
    D = np.zeros((n, 3))
    V = np.zeros((n, 3, 3))   # V[i, c, r] in memory = Fortran V(r, c, i)
    L = load_library()
    rc = L.epic_diagonalise(n, _dp(a), _dp(D), _dp(V))
    if rc != 0:
        raise EpicError(rc, L.epic_diagonalise_last_error().decode())
    return D, np.ascontiguousarray(np.transpose(V, (0, 2, 1)))


def next_write_time(freq, t):
    """get_next_write_time (utils.f90:48-54)."""
    return load_library().epic_next_write_time(float(freq), float(t))


def nc_file_info(path):
    """(file_type, last time, number of records, (nx, ny, nz, lower, extent))
    of a file (netcdf_reader.f90 get_file_type / get_time / get_num_steps)."""
    ft = C.create_string_buffer(16)
    t, n, g = C.c_double(0), C.c_int64(0), Grid()
    _nc_chk(load_library().epic_nc_file_info(str(path).encode(), ft, C.byref(t), C.byref(n), C.byref(g)))
    return ft.value.decode(), t.value, n.value, (g.nx, g.ny, g.nz, tuple(g.lower), tuple(g.extent))


def nc_read_fields(path, grid, step=-1):
    """Field file to halo-layout host arrays (epic_nc_read_fields): returns
    (dict vortg (3,H), tbuoyg, humg, t)."""
    nx, ny, nz = grid[0], grid[1], grid[2]
    g = Grid(nx, ny, nz, (C.c_double * 3)(*grid[3]), (C.c_double * 3)(*grid[4]))
    H = (nx + 3, ny + 3, nz + 3)
    vor, tb, hm, t = np.zeros((3,) + H), np.zeros(H), np.zeros(H), C.c_double(0)
    _nc_chk(load_library().epic_nc_read_fields(str(path).encode(), int(step), C.byref(g), _dp(vor), _dp(tb), _dp(hm),
                                                C.byref(t)))
    return {"vortg": vor, "tbuoyg": tb, "humg": hm}, t.value


def _nc_chk(rc):
    if rc != 0:
        raise EpicError(rc, load_library().epic_nc_last_error().decode())


def _config(nx, ny, nz, lower, extent, glat=0.0, lambda_c=0.0, q0=0.0, f_cor=(0.0, 0.0, 0.0),
            zeta_zero=(False, False)):
    g = Grid(nx, ny, nz, (C.c_double * 3)(*lower), (C.c_double * 3)(*extent))
    p = Phys(glat, lambda_c, q0, (C.c_double * 3)(*f_cor))
    f = Flags((C.c_int * 2)(*[int(bool(z)) for z in zeta_zero]), 4, 0.0)
    return g, p, f


def nc_write_parcels(path, grid, parcels, t, phys=None):
    """Parcel file from host arrays (epic_nc_write_parcels); grid = (nx, ny,
    nz, lower, extent); parcels = dict of position (n,3), vorticity (n,3),
    B (n,5), volume, buoyancy, humidity (n,)."""
    g, p, f = _config(*grid, **(phys or {}))
    arrs = [np.ascontiguousarray(parcels[k], dtype=np.float64)
            for k in ("position", "vorticity", "B", "volume", "buoyancy", "humidity")]
    n = arrs[0].shape[0]
    _nc_chk(load_library().epic_nc_write_parcels(str(path).encode(), C.byref(g), C.byref(p), C.byref(f), n,
                                                 *[_dp(a) for a in arrs], t))


def nc_read_parcels(path):
    """Parcel file to host arrays (epic_nc_read_parcels); returns (parcels,
    (nx, ny, nz, lower, extent), t)."""
    L = load_library()
    n, t, g = C.c_int64(0), C.c_double(0), Grid()
    _nc_chk(L.epic_nc_read_parcels(str(path).encode(), 0, C.byref(n), C.byref(g), None, None, None, None, None,
                                   None, C.byref(t)))
    m = n.value
    out = {"position": np.zeros((m, 3)), "vorticity": np.zeros((m, 3)), "B": np.zeros((m, 5)),
           "volume": np.zeros(m), "buoyancy": np.zeros(m), "humidity": np.zeros(m)}
    _nc_chk(L.epic_nc_read_parcels(str(path).encode(), m, C.byref(n), None,
                                   *[_dp(out[k]) for k in ("position", "vorticity", "B", "volume", "buoyancy",
                                                           "humidity")], C.byref(t)))
    return out, (g.nx, g.ny, g.nz, tuple(g.lower), tuple(g.extent)), t.value


def nc_write_fields(path, grid, fields, t, phys=None):
    """Field file from halo-layout host arrays (epic_nc_write_fields);
    fields = dict velog (3,H), vtend (3,H), vortg (3,H), tbuoyg, dbuoyg, humg,
    volg, nparg, nsparg (H) with H = (nx+3)(ny+3)(nz+3)."""
    g, p, f = _config(*grid, **(phys or {}))
    names = ("velog", "vtend", "vortg", "tbuoyg", "dbuoyg", "humg", "volg", "nparg", "nsparg")
    arrs = [np.ascontiguousarray(fields[k], dtype=np.float64) for k in names]
    _nc_chk(load_library().epic_nc_write_fields(str(path).encode(), C.byref(g), C.byref(p), C.byref(f), t,
                                                *[_dp(a) for a in arrs]))


class Batch:
    """Many members advanced by one batched LS-RK step (one launch per kernel
    over all members, one captured HIP graph per step)."""

    def __init__(self, members):
        L = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = C.c_void_p()
        rc = L.epic_batch_create(arr, len(self.members), C.byref(self._h))
        if rc != 0:
            raise EpicError(rc, "epic_batch_create failed")
        _live_batches.add(self)

    def step(self, t, dt, t_initial=None):
        """One step of every member; t, dt, t_initial are per-member sequences.
        Returns (new t array, status array)."""
        nm = len(self.members)
        tt = np.ascontiguousarray(t, dtype=np.float64).copy()
        dd = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (nm,)))
        ti = np.zeros(nm) if t_initial is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(t_initial, dtype=np.float64), (nm,)))
        st = np.zeros(nm, dtype=np.int32)
        L = load_library()
        rc = L.epic_batch_step(self._h, _dp(tt), _dp(dd), _dp(ti), st.ctypes.data_as(C.POINTER(C.c_int)))
        if rc != 0:
            raise EpicError(rc, (L.epic_batch_last_error(self._h) or b"").decode())
        return tt, st

    def prepare(self):
        """Builds and captures the step graph without running it."""
        L = load_library()
        rc = L.epic_batch_prepare(self._h)
        if rc != 0:
            raise EpicError(rc, (L.epic_batch_last_error(self._h) or b"").decode())

    def launch(self, t, dt, t_initial=None):
        """Queues one step of every member and returns at once; wait()
        completes it.  The members must not be used in between."""
        nm = len(self.members)
        self._t = np.ascontiguousarray(t, dtype=np.float64).copy()
        dd = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (nm,)))
        ti = np.zeros(nm) if t_initial is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(t_initial, dtype=np.float64), (nm,)))
        L = load_library()
        rc = L.epic_batch_launch(self._h, _dp(self._t), _dp(dd), _dp(ti))
        if rc != 0:
            raise EpicError(rc, (L.epic_batch_last_error(self._h) or b"").decode())

    def wait(self):
        """Completes the launched step: (new t array, status array)."""
        st = np.zeros(len(self.members), dtype=np.int32)
        L = load_library()
        rc = L.epic_batch_wait(self._h, _dp(self._t), st.ctypes.data_as(C.POINTER(C.c_int)))
        if rc != 0:
            raise EpicError(rc, (L.epic_batch_last_error(self._h) or b"").decode())
        return self._t, st

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load_library().epic_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
