"""Writes tests/golden/example_crop.npz: a 160 x 240 crop of the image pair
the reference's example runs on (data/example/1.jpg and 2.jpg,
examples/simple.py:8-9), decoded with PIL and stored in the B, G, R channel
order that example's ``cv2.imread`` gives.

The crop is data only (pixels of the reference's own example images), so
that a parity test runs the field on real photographs where the reference
tree is absent.  Run in a container that has the reference:

    python tests/golden/make_example_crop.py
"""
from __future__ import annotations

import os

import numpy as np
from PIL import Image

SRC = "/root/reference/data/example"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_crop.npz")
Y0, X0, H, W = 100, 1110, 160, 240


def main():
    imgs = []
    for name in ("1.jpg", "2.jpg"):
        with Image.open(os.path.join(SRC, name)) as im:
            rgb = np.asarray(im.convert("RGB"))
        imgs.append(np.ascontiguousarray(rgb[Y0:Y0 + H, X0:X0 + W, ::-1]))
    np.savez_compressed(OUT, a=imgs[0], b=imgs[1], origin=np.array([Y0, X0], np.int64))
    print(OUT, imgs[0].shape, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
