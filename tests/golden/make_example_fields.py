"""Write example_fields.npz (run from the repo root where the reference tree
exists):

    python tests/golden/make_example_fields.py

The coefficient data files the reference ships for its examples
(examples/ExampleDataFiles/) as arrays, since the GPU tests cannot read the
reference tree:

* ``gmm3_*``   gmm3_120_sha.tab: the GMM-3 gravity model of Mars, 4pi-
  normalised, degree 120; its header's reference radius (km) and GM
  (km^3/s^2).  examples/python/GravMag/TestGrav.py feeds it to
  MakeGravGridDH.
* ``mars90``   FSU_mars90.sh: Mars topography, degree 90.
* ``srtm300``  srtmp300.msl: Earth topography, degree 300.

Each data line holds l, m, C and (except where m = 0 in FSU_mars90.sh) S;
absent coefficients are zero.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = "/root/reference/examples/ExampleDataFiles"


def read_coefficients(path, skip, lmax):
    """cilm[2, lmax+1, lmax+1] from the data lines after `skip` header lines."""
    cilm = np.zeros((2, lmax + 1, lmax + 1))
    with open(path) as fh:
        lines = fh.read().splitlines()[skip:]
    for ln in lines:
        tok = [t for t in re.split(r"[,\s]+", ln.strip()) if t]
        if len(tok) < 3:
            continue
        l, m = int(tok[0]), int(tok[1])
        if l > lmax:
            continue
        cilm[0, l, m] = float(tok[2])
        if len(tok) > 3:
            cilm[1, l, m] = float(tok[3])
    return cilm


def main():
    with open(os.path.join(DATA, "gmm3_120_sha.tab")) as fh:
        head = [t for t in re.split(r"[,\s]+", fh.readline().strip()) if t]
    r0_km, gm_km3, lmax = float(head[0]), float(head[1]), int(head[3])
    gmm3 = read_coefficients(os.path.join(DATA, "gmm3_120_sha.tab"), 1, lmax)
    mars90 = read_coefficients(os.path.join(DATA, "FSU_mars90.sh"), 2, 90)
    srtm300 = read_coefficients(os.path.join(DATA, "srtmp300.msl"), 0, 300)
    np.savez_compressed(os.path.join(HERE, "example_fields.npz"), gmm3_cilm=gmm3, gmm3_r0_km=r0_km,
                        gmm3_gm_km3=gm_km3, mars90=mars90, srtm300=srtm300)


if __name__ == "__main__":
    main()
