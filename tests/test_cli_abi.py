"""The boundary without a GPU: the C ABI's exports and the drop-in CLI's
argument handling.

Every function that include/siting.h declares must be exported by
libsiting_amd.so.  The CLI must reject bad invocations exactly as the
reference does (siting.cpp:22-25, 663-695, 726-729): same stdout text, same
exit status.  All the cases below fail before any GPU work, so they run
here; when the reference build is present (oracle/_ref/seqsite, compiled from
the reference's own source) its output is the expected one, otherwise the
messages are spelled out.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "siting.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = set(re.findall(r"\b(siting_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_stage_functions():
    names = declared_functions()
    for fn in ("siting_calc_vix", "siting_find_max", "siting_calc_sheds", "siting_site_it",
               "siting_job_run", "siting_batch_run", "siting_server_submit"):
        assert fn in names


def test_abi_exports_every_declared_symbol(sa):
    L = sa.lib()
    missing = [fn for fn in declared_functions() if not hasattr(L, fn)]
    assert not missing, f"declared in siting.h but not exported: {missing}"


def test_status_strings(sa):
    L = sa.lib()
    import ctypes
    L.siting_status_str.restype = ctypes.c_char_p
    L.siting_status_str.argtypes = [ctypes.c_int]
    assert L.siting_status_str(0) == b"ok"
    assert L.siting_status_str(3) == b"The last block is too small for nwantedperblock."
    for code in range(0, 9):
        assert L.siting_status_str(code)


# (argv after the program name, expected stdout, expected exit status)
def _cases(tmp):
    short = os.path.join(tmp, "short.bin")
    np.zeros(99 * 100, np.uint16).tofile(short)  # one row short of 100^2
    missing = os.path.join(tmp, "does_not_exist.bin")
    out = os.path.join(tmp, "obs.txt")
    usage = ("ERROR: SITE requires 9 arguments: nrows, roi, oht/tht, ntests, blocksize, "
             "nwanted, intervis, infile, outfile\n")

    def a(n=100, roi=10, h=10, t=10, bw=10, nw=100, iv=0, inp=missing):
        return [str(n), str(roi), str(h), str(t), str(bw), str(nw), str(iv), inp, out]

    return [
        ([], "argc=1\n" + usage, 1),
        (a()[:8], "argc=9\n" + usage, 1),
        (a() + ["extra"], "argc=11\n" + usage, 1),
        (a(n=0), "ERROR: Unreasonable value for nrows\n", 1),
        (a(n=20001), "ERROR: Unreasonable value for nrows\n", 1),
        (a(roi=0), "ERROR: Unreasonable value for roi.\n", 1),
        (a(roi=10001), "ERROR: Unreasonable value for roi.\n", 1),
        (a(h=-1), "ERROR: Unreasonable value for tht.\n", 1),
        (a(h=1000001), "ERROR: Unreasonable value for tht.\n", 1),
        (a(t=0), "ERROR: Unreasonable value for ntests.\n", 1),
        (a(t=1001), "ERROR: Unreasonable value for ntests.\n", 1),
        (a(bw=9), "ERROR: Unreasonable value for blocksize0.\n", 1),
        (a(bw=2001), "ERROR: Unreasonable value for blocksize0.\n", 1),
        (a(nw=99), "ERROR: Unreasonable value for nwanted0.\n", 1),
        (a(nw=2000001), "ERROR: Unreasonable value for nwanted0.\n", 1),
        (a(iv=2), "ERROR: Unreasonable value for intervis.\n", 1),
        (a(iv=-1), "ERROR: Unreasonable value for intervis.\n", 1),
        # the first failing check wins (nrows before roi)
        (a(n=0, roi=0), "ERROR: Unreasonable value for nrows\n", 1),
        # non-numeric arguments read as 0 (atoi)
        (a(n="abc"), "ERROR: Unreasonable value for nrows\n", 1),
        # 10 blocks of 10, nwpb 20000 > 10^2 points in the last block
        (a(nw=2000000), "ERROR: The last block is too small for nwantedperblock.\n", 1),
        (a(inp=missing), "ERROR: Input failed\n", 1),
        (a(inp=short), "ERROR: Input failed\n", 1),
    ]


def test_cli_rejects_like_the_reference(sa, tmp_path):
    assert os.path.exists(sa.CLI_PATH), "hipsite is not built"
    import oracle
    seq = oracle.REF_SEQSITE if os.path.exists(oracle.REF_SEQSITE) else None
    for argv, want, rc in _cases(str(tmp_path)):
        got = subprocess.run([sa.CLI_PATH] + argv, capture_output=True, text=True,
                             timeout=60)
        if seq:
            ref = subprocess.run([seq] + argv, capture_output=True, text=True, timeout=60)
            assert (got.stdout, got.returncode) == (ref.stdout, ref.returncode), argv
        assert got.stdout == want, argv
        assert got.returncode == rc, argv


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_check_cli_agrees_with_oracle(sa, orc, seed):
    rng = np.random.default_rng(seed)
    for _ in range(400):
        vals = dict(nrows=int(rng.integers(-5, 20010)), roi=int(rng.integers(-2, 10010)),
                    ht=int(rng.integers(-3, 1000010)), ntests=int(rng.integers(-1, 1005)),
                    bw=int(rng.integers(0, 2010)), nw=int(rng.integers(90, 2000010)),
                    iv=int(rng.integers(-1, 3)))
        p = sa.make_params(vals["nrows"], vals["roi"], vals["ht"], vals["ntests"], vals["bw"],
                           vals["nw"], vals["iv"])
        which, _ = sa.check_cli(p)
        want = orc.lib().so_check_args(vals["nrows"], vals["roi"], vals["ht"], vals["ntests"],
                                       vals["bw"], vals["nw"], vals["iv"])
        assert which == want, vals


def test_dem_round_trip_and_short_file(sa, tmp_path):
    dem = (np.arange(37 * 37, dtype=np.uint32) * 2654435761 % 65536).astype(np.uint16)
    path = str(tmp_path / "d.bin")
    sa.dem_write(path, 37, dem)
    assert os.path.getsize(path) == 37 * 37 * 2
    assert np.array_equal(sa.dem_read(path, 37), dem)
    # extra bytes are ignored (siting.cpp:727 reads exactly nrows^2 values)
    with open(path, "ab") as f:
        f.write(b"\x01\x02\x03")
    assert np.array_equal(sa.dem_read(path, 37), dem)
    with pytest.raises(sa.SitingError):
        sa.dem_read(path, 38)


def test_library_roi_bound(sa):
    """The library accepts roi up to the CLI's bound of 10000
    (siting.cpp:680-686) and rejects larger ones with SITING_EINVAL."""
    d = sa.derive(sa.make_params(3000, 10000, 100, 64, 16, 10000, 0))
    assert d.nr == 2 * 10000 + 1
    with pytest.raises(sa.SitingError) as err:
        sa.derive(sa.make_params(3000, 10001, 100, 64, 16, 10000, 0))
    assert err.value.status == sa.EINVAL
