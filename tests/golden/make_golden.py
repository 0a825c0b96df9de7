"""Writes the golden fixtures under tests/golden/ from the reference build.

Run in a container that has the reference sources (oracle/Makefile `ref`
target compiles them into oracle/_ref/):

    python tests/golden/make_golden.py

Fixtures (data only: inputs, expected outputs, digests):
  dem1000.u16.gz        the reference's sample DEM (dem.bin, README.md:16),
                        gzip-compressed, sha256 checked on load
  readme_example.json   seqsite on it with the README example parameters
                        (README.md:38): observers file text, nusedsheds and
                        coverage as printed
  stages_small.npz      full stage outputs of the reference's stage
                        functions on tiny generated DEMs (vix, vix after
                        FINDMAX, obs, sheds, selected)
  stages_digest.json    sha256 digests of the same outputs on larger
                        generated DEMs (bw 8 and intervis 1 included), and of
                        the generated DEMs themselves
"""
import ctypes
import gzip
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import siting_amd as sa  # noqa: E402

REF_DEM = "/root/reference/dem.bin"

# (n, roi, height, ntests, bw, nwanted0, intervis, rough_h, hmax, seed)
SMALL = [
    (48, 6, 100, 16, 8, 64, 0, 0.8, 20000, 101),
    (60, 9, 50, 40, 10, 100, 1, 0.6, 30000, 102),
    (75, 12, 100, 70, 7, 300, 1, 0.7, 20000, 103),
]
MEDIUM = [
    (300, 30, 100, 50, 10, 900, 0, 0.8, 20000, 201),
    (256, 25, 50, 100, 8, 512, 1, 0.6, 30000, 202),
    (333, 40, 100, 64, 16, 1000, 0, 0.9, 20000, 203),
    (512, 50, 100, 32, 8, 2048, 1, 0.5, 20000, 204),
]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def dem_for(n, H, hmax, seed):
    return sa.dem_generate(sa.TerrainSpec(nrows=n, rough_h=H, hmin=0, hmax=hmax,
                                          nhills=8, hill_peak=5000.0, hill_sigma=80.0,
                                          seed=seed))


def ref_stages(n, dem, roi, ht, nt, bw, nw0, iv):
    """The reference's own stage functions (oracle/_ref/libsiting_ref.so)."""
    R = oracle.ref_lib()
    rc, d = oracle.derive(n, bw, nw0)
    assert rc == 0
    vix = np.empty(n * n, np.uint8)
    R.ref_calc_vix(n, dem.ctypes.data, roi, ht, ht, nt, vix.ctypes.data)
    vix0 = vix.copy()
    obs = np.empty(d.nwanted, np.int32)
    R.ref_find_max(n, vix.ctypes.data, d.blocksize, d.nwanted, d.nblocks, d.nwpb,
                   obs.ctypes.data)
    nr = 2 * roi + 1
    nwpr = (nr + 63) // 64
    sheds = np.empty(d.nwanted * nr * nwpr, np.uint64)
    R.ref_calc_sheds(n, dem.ctypes.data, roi, ht, ht, d.nwanted, obs.ctypes.data,
                     sheds.ctypes.data)
    sel = np.empty(d.nwanted, np.uint8)
    sys.stdout.flush()
    R.ref_site_it(n, roi, iv, d.nwanted, obs.ctypes.data, sheds.ctypes.data,
                  sel.ctypes.data)
    print(flush=True)
    return vix0, vix, obs, sheds, sel


def main():
    oracle.build(ref=True)
    raw = open(REF_DEM, "rb").read()
    with gzip.open(os.path.join(HERE, "dem1000.u16.gz"), "wb", compresslevel=9) as f:
        f.write(raw)

    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "observers.txt")
        r = subprocess.run([oracle.REF_SEQSITE, "1000", "50", "100", "50", "10",
                            "10000", "0", REF_DEM, out], capture_output=True,
                           text=True, check=True)
        toks = dict(t.split(":") for t in r.stdout.split())
        readme = {"args": [1000, 50, 100, 50, 10, 10000, 0],
                  "dem_sha256": hashlib.sha256(raw).hexdigest(),
                  "nusedsheds": int(toks["nusedsheds"]),
                  "coverage_printed": toks["coverage"],
                  "observers": open(out).read()}
    with open(os.path.join(HERE, "readme_example.json"), "w") as f:
        json.dump(readme, f, indent=1)

    arrays = {}
    for i, c in enumerate(SMALL):
        n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = c
        dem = dem_for(n, H, hmax, seed)
        vix0, vix1, obs, sheds, sel = ref_stages(n, dem, roi, ht, nt, bw, nw0, iv)
        arrays[f"c{i}_case"] = np.array(c, dtype=np.float64)
        arrays[f"c{i}_dem"] = dem
        arrays[f"c{i}_vix"] = vix0
        arrays[f"c{i}_vix_after"] = vix1
        arrays[f"c{i}_obs"] = obs
        arrays[f"c{i}_sheds"] = sheds
        arrays[f"c{i}_selected"] = sel
    np.savez_compressed(os.path.join(HERE, "stages_small.npz"), **arrays)

    digests = []
    for c in MEDIUM:
        n, roi, ht, nt, bw, nw0, iv, H, hmax, seed = c
        dem = dem_for(n, H, hmax, seed)
        vix0, vix1, obs, sheds, sel = ref_stages(n, dem, roi, ht, nt, bw, nw0, iv)
        digests.append({"case": list(c), "dem": sha(dem), "vix": sha(vix0),
                        "vix_after": sha(vix1), "obs": sha(obs), "sheds": sha(sheds),
                        "selected": sha(sel), "nused": int(sel.sum())})
    with open(os.path.join(HERE, "stages_digest.json"), "w") as f:
        json.dump(digests, f, indent=1)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
