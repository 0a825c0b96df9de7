"""CPU oracle for the siting pipeline -- TEST INFRASTRUCTURE ONLY.

ctypes view of ``oracle/liboracle.so``, the project's C restatement of the
reference's sequential path (siting.cpp built -DSEQ), plus, when it has been
built in this tree, ``oracle/_ref/libsiting_ref.so`` (the reference's own
siting.cpp compiled from where it lies by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, and only as the checker / CPU baseline; the product (siting_amd)
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libsiting_ref.so")
REF_LIB_WRAPV = os.path.join(REF_DIR, "libsiting_ref_wrapv.so")
REF_SEQSITE = os.path.join(REF_DIR, "seqsite")
REF_OMPSITE = os.path.join(REF_DIR, "ompsite")

_VP = ctypes.c_void_p
_I = ctypes.c_int


class SoDerived(ctypes.Structure):
    _fields_ = [("blocksize", ctypes.c_float), ("nblocks", _I), ("nwpb", _I),
                ("nwanted", _I), ("lastsize", _I)]


class SoSiteStats(ctypes.Structure):
    _fields_ = [("nused", _I), ("niter", _I), ("cumarea", ctypes.c_int64),
                ("sum_u", ctypes.c_int64), ("coverage", ctypes.c_double)]


_lib = None
_ref = {}


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and _ref/ when ref=True and the reference
    sources are present)."""
    target = "ref" if ref else "all"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.so_lcg.restype = ctypes.c_int32
        L.so_lcg.argtypes = [ctypes.c_int32]
        L.so_derive.argtypes = [_I, _I, _I, ctypes.POINTER(SoDerived)]
        L.so_check_args.argtypes = [_I] * 7
        L.so_calc_vix.argtypes = [_I, _VP, _I, _I, _I, _I, _VP, _I, _VP]
        L.so_test_one_target.argtypes = [_I, _VP, _I, _I, _I, _I, _I, _I]
        L.so_find_max.argtypes = [_I, _VP, ctypes.c_float, _I, _I, _VP, _I]
        L.so_calc_sheds.argtypes = [_I, _VP, _I, _I, _I, _I, _VP, _VP, _I, _VP]
        L.so_site.argtypes = [_I, _I, _I, _I, _VP, _VP, _VP, _VP,
                              ctypes.POINTER(SoSiteStats), _I]
        L.so_union_area.argtypes = [_I, _I, _I, _I, _VP, _VP]
        L.so_calc_union.argtypes = [_I, _I, _I, _I, _VP, _VP]
        L.so_run.argtypes = [_I, _VP, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP,
                             ctypes.POINTER(SoSiteStats), _VP, _I]
        _lib = L
    return _lib


def ref_available(wrapv: bool = False) -> bool:
    return os.path.exists(REF_LIB_WRAPV if wrapv else REF_LIB)


def ref_lib(wrapv: bool = False) -> ctypes.CDLL:
    key = "wrapv" if wrapv else "plain"
    if key not in _ref:
        L = ctypes.CDLL(REF_LIB_WRAPV if wrapv else REF_LIB)
        L.ref_calc_vix.argtypes = [_I, _VP, _I, _I, _I, _I, _VP]
        L.ref_find_max.argtypes = [_I, _VP, ctypes.c_float, _I, _I, _I, _VP]
        L.ref_calc_sheds.argtypes = [_I, _VP, _I, _I, _I, _I, _VP, _VP]
        L.ref_site_it.argtypes = [_I, _I, _I, _I, _VP, _VP, _VP]
        L.ref_test_one_target.argtypes = [_I, _VP, _I, _I, _I, _I, _I, _I]
        _ref[key] = L
    return _ref[key]


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def derive(nrows: int, bw: int, nwanted0: int) -> tuple[int, SoDerived]:
    d = SoDerived()
    rc = lib().so_derive(nrows, bw, nwanted0, ctypes.byref(d))
    return rc, d


def lcg(seed: int) -> int:
    return lib().so_lcg(seed)


def calc_vix(nrows, elevs, roi, oht, tht, ntests, nthreads=1, count=False):
    vix = np.empty(nrows * nrows, np.uint8)
    c = ctypes.c_int64(0)
    lib().so_calc_vix(nrows, _p(elevs), roi, oht, tht, ntests, _p(vix), nthreads,
                      ctypes.byref(c) if count else None)
    return (vix, c.value) if count else vix


def find_max(nrows, vix, blocksize, nblocks, nwpb, nthreads=1):
    obs = np.empty(nblocks * nblocks * nwpb, np.int32)
    lib().so_find_max(nrows, _p(vix), blocksize, nblocks, nwpb, _p(obs), nthreads)
    return obs


def calc_sheds(nrows, elevs, roi, oht, tht, obs, nthreads=1, count=False):
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    o = np.ascontiguousarray(obs, dtype=np.int32)
    sheds = np.empty(o.size * nr * nwpr, np.uint64)
    c = ctypes.c_int64(0)
    lib().so_calc_sheds(nrows, _p(elevs), roi, oht, tht, o.size, _p(o), _p(sheds),
                        nthreads, ctypes.byref(c) if count else None)
    return (sheds, c.value) if count else sheds


def site(nrows, roi, intervis, obs, sheds, nthreads=1):
    o = np.ascontiguousarray(obs, dtype=np.int32)
    sel = np.empty(o.size, np.uint8)
    order = np.empty(o.size, np.int32)
    st = SoSiteStats()
    lib().so_site(nrows, roi, intervis, o.size, _p(o), _p(sheds), _p(sel), _p(order),
                  ctypes.byref(st), nthreads)
    return sel, order[: st.nused].copy(), st


def run(nrows, elevs, roi, ht, ntests, bw, nwanted0, intervis, nthreads=1):
    """Whole pipeline; returns a dict of every stage output."""
    rc, d = derive(nrows, bw, nwanted0)
    if rc:
        raise ValueError(f"derive failed ({rc})")
    vix = calc_vix(nrows, elevs, roi, ht, ht, ntests, nthreads)
    vix_before = vix.copy()
    obs = find_max(nrows, vix, d.blocksize, d.nblocks, d.nwpb, nthreads)
    sheds = calc_sheds(nrows, elevs, roi, ht, ht, obs, nthreads)
    sel, order, st = site(nrows, roi, intervis, obs, sheds, nthreads)
    return {"derived": d, "vix": vix_before, "vix_after": vix, "obs": obs,
            "sheds": sheds, "selected": sel, "order": order, "stats": st}
