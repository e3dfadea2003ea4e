"""Writes tv_images.npz: the images the reference's regularization-path check
uses (scripts/test_reg_path.py:11-26 reads benchmarks/images/*.png), after that
script's preprocessing: the mean over the colour channels (alpha included, as
matplotlib's imread returns it) divided by its maximum, every second row and
column, then centred and scaled to unit standard deviation.

Run from the repository root with the reference checked out at
/root/reference; the output is committed, so tests never read the reference."""
import os

import numpy as np
from PIL import Image

SRC = "/root/reference/benchmarks/images"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tv_images.npz")


def load(name, step):
    a = np.asarray(Image.open(os.path.join(SRC, name)), dtype=np.float64) / 255.0
    X = (a.mean(axis=2) / a.max())[::step, ::step]
    X = X - X.mean()
    sd = X.std()
    return X / sd if sd > 0 else X


if __name__ == "__main__":
    np.savez_compressed(OUT, sanity=load("sanity.png", 1), truffles_small=load("truffles-small.png", 2),
                        branches_small=load("branches-small.png", 2))
    with np.load(OUT) as z:
        for k in z.files:
            print(k, z[k].shape)
