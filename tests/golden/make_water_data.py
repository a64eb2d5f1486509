"""Copy the first frames of the reference's water training system into
tests/golden/water_data_0 (type.raw, type_map.raw, set.000/{box,coord,energy,
force}.npy), the data the training-loop tests run on.

Source: source/tests/pt/model/water/data/data_0 of the reference (192 atoms,
O/H, 80 frames; the input its se_e2_a training test uses,
source/tests/pt/water/se_e2_a.json).  Arrays are loaded with
allow_pickle=False and written unchanged (float32), first NFRAMES frames.
Run in the build container: python tests/golden/make_water_data.py
"""

import os
import shutil

import numpy as np

SRC = "/root/reference/source/tests/pt/model/water/data/data_0"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water_data_0")
NFRAMES = 6


def main():
    os.makedirs(os.path.join(DST, "set.000"), exist_ok=True)
    for f in ("type.raw", "type_map.raw"):
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, f))
    for key in ("box", "coord", "energy", "force"):
        a = np.load(os.path.join(SRC, "set.000", key + ".npy"), allow_pickle=False)
        np.save(os.path.join(DST, "set.000", key + ".npy"), a[:NFRAMES])


if __name__ == "__main__":
    main()
