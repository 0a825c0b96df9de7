"""MI355X-native multiple-observer siting (VIX -> FINDMAX -> VIEWSHED -> SITE).

Python view of the C ABI in ``include/siting.h`` (``lib/libsiting_amd.so``).
The four stage functions keep the reference's names, argument order and
meaning (siting.cpp:192, 235, 349, 501); they take and return numpy arrays
and raise :class:`SitingError` where the reference would ``die()``.

There is no CPU fallback: importing works without a GPU, but every compute
call needs the HIP library and a device, and fails loudly otherwise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsiting_amd.so")
# the same library with the greedy loop's phase counters (siting_amd.phases)
PROF_LIB_PATH = os.path.join(_HERE, "lib", "libsiting_amd_prof.so")
CLI_PATH = os.path.join(_HERE, "bin", "hipsite")

OK, EINVAL, ERANGE, ESMALLBLOCK, ENOMEM, EDEVICE, EIO, EBUSY, ENOTFOUND = range(9)

BUF_ELEVS, BUF_VIX, BUF_OBS, BUF_SHEDS, BUF_SELECTED, BUF_ORDER = range(6)


class SitingError(RuntimeError):
    """A non-OK siting_status from the library."""

    def __init__(self, status: int, message: str):
        super().__init__(f"siting status {status}: {message}")
        self.status = status
        self.message = message


class Params(ctypes.Structure):
    _fields_ = [
        ("nrows", ctypes.c_int),
        ("roi", ctypes.c_int),
        ("oht", ctypes.c_int),
        ("tht", ctypes.c_int),
        ("ntests", ctypes.c_int),
        ("bw", ctypes.c_int),
        ("nwanted0", ctypes.c_int),
        ("intervis", ctypes.c_int),
    ]


class Derived(ctypes.Structure):
    _fields_ = [
        ("blocksize", ctypes.c_float),
        ("nblocks", ctypes.c_int),
        ("nwpb", ctypes.c_int),
        ("nwanted", ctypes.c_int),
        ("lastsize", ctypes.c_int),
        ("nr", ctypes.c_int),
        ("nwpr", ctypes.c_int),
        ("cumnwpr", ctypes.c_int),
        ("shed_words", ctypes.c_int64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("nused", ctypes.c_int),
        ("niter", ctypes.c_int),
        ("cumarea", ctypes.c_int64),
        ("sum_u", ctypes.c_int64),
        ("coverage", ctypes.c_double),
        ("ms_vix", ctypes.c_float),
        ("ms_findmax", ctypes.c_float),
        ("ms_viewshed", ctypes.c_float),
        ("ms_site", ctypes.c_float),
        ("rounds", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(ctypes.Structure):
    """Kernel shape choices of one job (siting.h, siting_tuning); 0 keeps a
    field's built-in rule."""
    _fields_ = [
        ("vix_halo", ctypes.c_int),
        ("viewshed_threads", ctypes.c_int),
        ("viewshed_mode", ctypes.c_int),
        ("site_q", ctypes.c_int),
        ("site_tau_shift", ctypes.c_int),
        ("site_walk_evals", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None

_VP = ctypes.c_void_p
_I = ctypes.c_int
_SIGS = {
    "siting_status_str": (ctypes.c_char_p, [_I]),
    "siting_derive": (_I, [ctypes.POINTER(Params), ctypes.POINTER(Derived)]),
    "siting_check_cli": (_I, [ctypes.POINTER(Params), ctypes.POINTER(_I),
                             ctypes.POINTER(ctypes.c_char_p)]),
    "siting_calc_vix": (_I, [_I, _VP, _I, _I, _I, _I, _VP]),
    "siting_find_max": (_I, [_I, _VP, ctypes.c_float, _I, _I, _I, _VP]),
    "siting_calc_sheds": (_I, [_I, _VP, _I, _I, _I, _I, _VP, _VP]),
    "siting_site_it": (_I, [_I, _I, _I, _I, _VP, _VP, _VP, ctypes.POINTER(Stats)]),
    "siting_job_bytes": (_I, [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_size_t)]),
    "siting_job_create": (_I, [_I, ctypes.POINTER(Params), ctypes.POINTER(_VP)]),
    "siting_job_destroy": (None, [_VP]),
    "siting_job_upload_dem": (_I, [_VP, _VP]),
    "siting_job_vix": (_I, [_VP]),
    "siting_job_findmax": (_I, [_VP]),
    "siting_job_viewshed": (_I, [_VP]),
    "siting_job_site": (_I, [_VP]),
    "siting_job_run": (_I, [_VP]),
    "siting_job_sync": (_I, [_VP]),
    "siting_job_download": (_I, [_VP, _I, _VP, ctypes.c_size_t]),
    "siting_job_device_ptr": (_I, [_VP, _I, ctypes.POINTER(_VP),
                                   ctypes.POINTER(ctypes.c_size_t)]),
    "siting_job_stats": (_I, [_VP, ctypes.POINTER(Stats)]),
    "siting_job_derived": (_I, [_VP, ctypes.POINTER(Derived)]),
    "siting_dem_read": (_I, [ctypes.c_char_p, _I, _VP]),
    "siting_dem_write": (_I, [ctypes.c_char_p, _I, _VP]),
    "siting_dem_generate": (_I, [_I, ctypes.c_double, _I, _I, _I, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_uint64, _VP]),
    "siting_device_count": (_I, []),
    "siting_job_set_tuning": (_I, [_VP, _VP]),
    "siting_job_get_tuning": (_I, [_VP, _VP]),
    "siting_device_key": (_I, [_I, ctypes.c_char_p, ctypes.c_size_t]),
    "siting_tuning_lookup": (_I, [ctypes.c_char_p, ctypes.POINTER(Params), _VP]),
    "siting_tuning_cache_path": (ctypes.c_char_p, []),
    "siting_site_it_noalloc": (_I, [_I, _I, _I, _I, _VP, _VP, _VP, _VP,
                                    ctypes.POINTER(Stats), _VP]),
    "siting_batch_create": (_I, [_I, _I, _VP, _VP, ctypes.POINTER(_VP)]),
    "siting_batch_destroy": (None, [_VP]),
    "siting_batch_upload_dem": (_I, [_VP, _I, _VP]),
    "siting_batch_run": (_I, [_VP]),
    "siting_batch_sync": (_I, [_VP]),
    "siting_batch_elapsed_ms": (_I, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "siting_batch_site_ms": (_I, [_VP, ctypes.POINTER(ctypes.c_float)]),
    "siting_batch_site_times": (_I, [_VP, _VP, _I]),
    "siting_batch_job": (_I, [_VP, _I, ctypes.POINTER(_VP)]),
    "siting_server_create": (_I, [_I, _VP, ctypes.POINTER(_VP)]),
    "siting_server_destroy": (None, [_VP]),
    "siting_server_submit": (_I, [_VP, _VP, _VP, ctypes.POINTER(ctypes.c_uint64)]),
    "siting_server_info": (_I, [_VP, ctypes.c_uint64, _VP]),
    "siting_server_wait": (_I, [_VP, ctypes.c_uint64, ctypes.c_double, _VP]),
    "siting_server_result": (_I, [_VP, ctypes.c_uint64, _VP, ctypes.c_size_t, _VP,
                                  ctypes.c_size_t]),
    "siting_server_cancel": (_I, [_VP, ctypes.c_uint64]),
    "siting_server_release": (_I, [_VP, ctypes.c_uint64]),
    "siting_server_load": (_I, [_VP, _VP]),
}


def lib() -> ctypes.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SitingError(EDEVICE, f"{LIB_PATH} is not built "
                              "(run make -C siting_amd/csrc)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def use_profiling_build() -> ctypes.CDLL:
    """Loads the profiling build in place of the product library; must come
    before anything else loads the library in this process."""
    global LIB_PATH
    if _lib is not None and LIB_PATH != PROF_LIB_PATH:
        raise SitingError(EINVAL, "the product library is already loaded")
    LIB_PATH = PROF_LIB_PATH
    handle = lib()
    handle.siting_phase_profile_read.restype = _I
    handle.siting_phase_profile_read.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    handle.siting_phase_trace_read.restype = _I
    handle.siting_phase_trace_read.argtypes = [_VP, ctypes.c_size_t]
    return handle


def _check(status: int) -> None:
    if status != OK:
        msg = lib().siting_status_str(status).decode()
        raise SitingError(status, msg)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    return lib().siting_device_count()


def make_params(nrows, roi, height=None, ntests=None, bw=None, nwanted0=None,
                intervis=0, oht=None, tht=None) -> Params:
    oht = height if oht is None else oht
    tht = height if tht is None else tht
    return Params(nrows, roi, oht, tht, ntests, bw, nwanted0, intervis)


def device_key(device: int = 0) -> str:
    """"<arch>:<CUs>" of a device, the tuning cache's key."""
    buf = ctypes.create_string_buffer(64)
    _check(lib().siting_device_key(device, buf, len(buf)))
    return buf.value.decode()


def tuning_lookup(key: str, params: Params):
    """The tuning cache's choices for these parameters on a device key, or
    None when the cache holds none."""
    t = Tuning()
    rc = lib().siting_tuning_lookup(key.encode(), ctypes.byref(params), ctypes.byref(t))
    if rc == ENOTFOUND:
        return None
    _check(rc)
    return t


def tuning_cache_path() -> str:
    return (lib().siting_tuning_cache_path() or b"").decode()


def derive(p: Params) -> Derived:
    """siting.cpp:689-702 (no CLI range checks)."""
    d = Derived()
    _check(lib().siting_derive(ctypes.byref(p), ctypes.byref(d)))
    return d


def check_cli(p: Params) -> tuple[int, str]:
    """siting.cpp:680-686: (0, "") or (index, die message)."""
    which = ctypes.c_int(0)
    msg = ctypes.c_char_p()
    lib().siting_check_cli(ctypes.byref(p), ctypes.byref(which), ctypes.byref(msg))
    return which.value, (msg.value or b"").decode()


# ---- stage functions (the reference's signatures) -------------------------

def calc_vix(nrows, elevs, roi, oht, tht, ntests) -> np.ndarray:
    """siting.cpp:192 calc_vix: uint8[nrows*nrows] visibility index."""
    e = np.ascontiguousarray(elevs, dtype=np.uint16).reshape(-1)
    if e.size != nrows * nrows:
        raise SitingError(EINVAL, "elevs must hold nrows*nrows values")
    vix = np.empty(nrows * nrows, np.uint8)
    _check(lib().siting_calc_vix(nrows, _ptr(e), roi, oht, tht, ntests, _ptr(vix)))
    return vix


def find_max(nrows, vix, blocksize, nwanted, nblocks, nwpb) -> np.ndarray:
    """siting.cpp:235 find_max: obs[nblocks^2 * nwpb]; vix is mutated in place
    (picked cells set to 0) as in the sequential reference."""
    if not (isinstance(vix, np.ndarray) and vix.dtype == np.uint8
            and vix.flags.c_contiguous and vix.size == nrows * nrows):
        raise SitingError(EINVAL, "vix must be a contiguous uint8 array of nrows^2")
    obs = np.empty(nblocks * nblocks * nwpb, np.int32)
    _check(lib().siting_find_max(nrows, _ptr(vix), blocksize, nwanted, nblocks,
                                 nwpb, _ptr(obs)))
    return obs


def calc_sheds(nrows, elevs, roi, oht, tht, nsheds, obs) -> np.ndarray:
    """siting.cpp:349 calc_sheds: uint64[nsheds * nr * nwpr] MSB-first."""
    e = np.ascontiguousarray(elevs, dtype=np.uint16).reshape(-1)
    o = np.ascontiguousarray(obs, dtype=np.int32)
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    if e.size != nrows * nrows or o.size < nsheds:
        raise SitingError(EINVAL, "bad elevs/obs size")
    sheds = np.empty(nsheds * nr * nwpr, np.uint64)
    _check(lib().siting_calc_sheds(nrows, _ptr(e), roi, oht, tht, nsheds, _ptr(o),
                                   _ptr(sheds)))
    return sheds


def site_it(nrows, roi, intervis, nsheds, obs, sheds):
    """siting.cpp:501 site_it: (selected uint8[nsheds], Stats)."""
    o = np.ascontiguousarray(obs, dtype=np.int32)
    s = np.ascontiguousarray(sheds, dtype=np.uint64)
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    if o.size < nsheds or s.size < nsheds * nr * nwpr:
        raise SitingError(EINVAL, "bad obs/sheds size")
    sel = np.empty(nsheds, np.uint8)
    st = Stats()
    _check(lib().siting_site_it(nrows, roi, intervis, nsheds, _ptr(o), _ptr(s),
                                _ptr(sel), ctypes.byref(st)))
    return sel, st


# ---- device-resident job ---------------------------------------------------

class Job:
    """One siting job whose arrays stay in HBM across the four stages."""

    def __init__(self, params: Params, device: int = 0):
        self.params = params
        h = ctypes.c_void_p()
        _check(lib().siting_job_create(device, ctypes.byref(params), ctypes.byref(h)))
        self._h = h
        d = Derived()
        _check(lib().siting_job_derived(self._h, ctypes.byref(d)))
        self.derived = d

    @staticmethod
    def bytes_needed(params: Params) -> int:
        b = ctypes.c_size_t()
        _check(lib().siting_job_bytes(ctypes.byref(params), ctypes.byref(b)))
        return b.value

    def close(self):
        if self._h:
            lib().siting_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_dem(self, elevs):
        e = np.ascontiguousarray(elevs, dtype=np.uint16).reshape(-1)
        if e.size != self.params.nrows ** 2:
            raise SitingError(EINVAL, "DEM size mismatch")
        _check(lib().siting_job_upload_dem(self._h, _ptr(e)))

    def vix(self):
        _check(lib().siting_job_vix(self._h))

    def findmax(self):
        _check(lib().siting_job_findmax(self._h))

    def viewshed(self):
        _check(lib().siting_job_viewshed(self._h))

    def site(self):
        _check(lib().siting_job_site(self._h))

    def run(self):
        _check(lib().siting_job_run(self._h))

    def sync(self):
        _check(lib().siting_job_sync(self._h))

    def stats(self) -> Stats:
        st = Stats()
        _check(lib().siting_job_stats(self._h, ctypes.byref(st)))
        return st

    def _download(self, which, dtype, count):
        out = np.empty(count, dtype)
        _check(lib().siting_job_download(self._h, which, _ptr(out), out.nbytes))
        return out

    def set_tuning(self, t: "Tuning | None"):
        _check(lib().siting_job_set_tuning(self._h, ctypes.byref(t) if t is not None else None))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().siting_job_get_tuning(self._h, ctypes.byref(t)))
        return t

    def download_into(self, which, out: np.ndarray):
        """Copies the first out.nbytes bytes of buffer `which` into out."""
        if not out.flags.c_contiguous:
            raise SitingError(EINVAL, "output array must be contiguous")
        _check(lib().siting_job_download(self._h, which, _ptr(out), out.nbytes))

    def download_vix(self):
        return self._download(BUF_VIX, np.uint8, self.params.nrows ** 2)

    def download_obs(self):
        return self._download(BUF_OBS, np.int32, self.derived.nwanted)

    def download_sheds(self):
        return self._download(BUF_SHEDS, np.uint64,
                              self.derived.nwanted * self.derived.shed_words)

    def download_selected(self):
        return self._download(BUF_SELECTED, np.uint8, self.derived.nwanted)

    def download_order(self, nused: int):
        return self._download(BUF_ORDER, np.int32, nused)


# ---- batches ---------------------------------------------------------------

class BatchOpts(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int), ("slot", ctypes.c_int),
                ("nslots", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("fixed_slots", ctypes.c_int)]


class _JobView(Job):
    """A job owned by a batch (not destroyed by this handle)."""

    def __init__(self, handle, params):
        self.params = params
        self._h = handle
        d = Derived()
        _check(lib().siting_job_derived(self._h, ctypes.byref(d)))
        self.derived = d

    def close(self):
        self._h = None


class Batch:
    """Many independent jobs run as one unit (see siting.h, batched jobs)."""

    def __init__(self, params_list, device=0, streams=0, slot=0, nslots=0, use_graph=False,
                 fixed_slots=False):
        arr = (Params * len(params_list))(*params_list)
        self.params = list(params_list)
        opts = BatchOpts(streams, slot, nslots, 1 if use_graph else 0,
                         1 if fixed_slots else 0)
        h = ctypes.c_void_p()
        _check(lib().siting_batch_create(device, len(params_list), ctypes.cast(arr, _VP),
                                         ctypes.byref(opts), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().siting_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_dem(self, i, elevs):
        e = np.ascontiguousarray(elevs, dtype=np.uint16).reshape(-1)
        if e.size != self.params[i].nrows ** 2:
            raise SitingError(EINVAL, "DEM size mismatch")
        _check(lib().siting_batch_upload_dem(self._h, i, _ptr(e)))

    def run(self):
        _check(lib().siting_batch_run(self._h))

    def sync(self):
        _check(lib().siting_batch_sync(self._h))

    def elapsed_ms(self) -> float:
        ms = ctypes.c_float()
        _check(lib().siting_batch_elapsed_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def site_ms(self) -> float:
        """Device time of the last run's SITE work-queue launch."""
        ms = ctypes.c_float()
        _check(lib().siting_batch_site_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def site_times(self) -> np.ndarray:
        """(njobs, 2) start and end of every job's SITE loop in the last run,
        ms from the first loop's start."""
        out = np.zeros(2 * len(self.params), np.float64)
        _check(lib().siting_batch_site_times(self._h, _ptr(out), len(self.params)))
        return out.reshape(-1, 2)

    def job(self, i) -> Job:
        h = ctypes.c_void_p()
        _check(lib().siting_batch_job(self._h, i, ctypes.byref(h)))
        return _JobView(h, self.params[i])


# ---- serving ---------------------------------------------------------------

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = range(4)


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_bytes", ctypes.c_size_t), ("max_active", ctypes.c_int),
                ("site_groups", ctypes.c_int), ("backfill_ms", ctypes.c_int)]


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("stats", Stats),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


class LoadInfo(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int), ("running", ctypes.c_int),
                ("groups_in_use", ctypes.c_int), ("hbm_bytes", ctypes.c_size_t)]


class Server:
    """Request loop on one GPU (see siting.h, serving)."""

    def __init__(self, device=0, hbm_bytes=0, max_active=0, site_groups=0, backfill_ms=0):
        opts = ServerOpts(hbm_bytes, max_active, site_groups, backfill_ms)
        h = ctypes.c_void_p()
        _check(lib().siting_server_create(device, ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self._params = {}

    def close(self):
        if self._h:
            lib().siting_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, params: Params, elevs) -> int:
        e = np.ascontiguousarray(elevs, dtype=np.uint16).reshape(-1)
        if e.size != params.nrows ** 2:
            raise SitingError(EINVAL, "DEM size mismatch")
        jid = ctypes.c_uint64()
        _check(lib().siting_server_submit(self._h, ctypes.byref(params), _ptr(e),
                                          ctypes.byref(jid)))
        self._params[jid.value] = params
        return jid.value

    def info(self, jid) -> JobInfo:
        info = JobInfo()
        _check(lib().siting_server_info(self._h, jid, ctypes.byref(info)))
        return info

    def wait(self, jid, timeout_s=-1.0) -> JobInfo:
        info = JobInfo()
        _check(lib().siting_server_wait(self._h, jid, timeout_s, ctypes.byref(info)))
        return info

    def result(self, jid):
        d = derive(self._params[jid])
        sel = np.empty(d.nwanted, np.uint8)
        obs = np.empty(d.nwanted, np.int32)
        _check(lib().siting_server_result(self._h, jid, _ptr(sel), sel.size, _ptr(obs),
                                          obs.size))
        return sel, obs

    def cancel(self, jid):
        _check(lib().siting_server_cancel(self._h, jid))

    def release(self, jid):
        _check(lib().siting_server_release(self._h, jid))
        self._params.pop(jid, None)

    def load(self) -> LoadInfo:
        li = LoadInfo()
        _check(lib().siting_server_load(self._h, ctypes.byref(li)))
        return li


# ---- terrain ---------------------------------------------------------------

@dataclass
class TerrainSpec:
    """Seeded synthetic DEM: diamond-square roughness H, linear rescale to
    [hmin, hmax], plus Gaussian hills."""
    nrows: int
    rough_h: float = 0.8
    hmin: int = 0
    hmax: int = 20000
    nhills: int = 8
    hill_peak: float = 5000.0
    hill_sigma: float = 80.0
    seed: int = 1


def dem_generate(spec: TerrainSpec) -> np.ndarray:
    out = np.empty(spec.nrows * spec.nrows, np.uint16)
    _check(lib().siting_dem_generate(spec.nrows, spec.rough_h, spec.hmin, spec.hmax,
                                     spec.nhills, spec.hill_peak, spec.hill_sigma,
                                     spec.seed & (2 ** 64 - 1), _ptr(out)))
    return out


def dem_read(path: str, nrows: int) -> np.ndarray:
    out = np.empty(nrows * nrows, np.uint16)
    _check(lib().siting_dem_read(path.encode(), nrows, _ptr(out)))
    return out


def dem_write(path: str, nrows: int, dem: np.ndarray) -> None:
    e = np.ascontiguousarray(dem, dtype=np.uint16).reshape(-1)
    _check(lib().siting_dem_write(path.encode(), nrows, _ptr(e)))
