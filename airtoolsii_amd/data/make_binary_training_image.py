"""Writes binary_training_image.txt, the training image of the 'binary'
phantom: the 0/1 matrix that testprobs/phantomgallery.m:677-2641 (getTI)
lists as the (row, column) positions of its ones, full(sparse(I, J, 1)).
The text file holds one image row per line ('0'/'1' characters), top row
first.  Run from the repository root with the reference checked out:

    python airtoolsii_amd/data/make_binary_training_image.py /root/reference
"""
import os
import sys

import numpy as np


def training_image(reference_root):
    src = open(os.path.join(reference_root, "testprobs", "phantomgallery.m")).read()
    start = src.index("function A = getTI")
    body = src[src.index("IJ = [", start) + len("IJ = ["):src.index("];", start)]
    ij = np.array([[int(t) for t in line.split()] for line in body.splitlines() if line.strip()])
    im = np.zeros((ij[:, 0].max(), ij[:, 1].max()), np.uint8)
    np.add.at(im, (ij[:, 0] - 1, ij[:, 1] - 1), 1)        # sparse() sums repeated pairs
    return im


if __name__ == "__main__":
    im = training_image(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    assert im.max() == 1, "the training image is binary"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "binary_training_image.txt")
    with open(out, "w") as f:
        f.write("\n".join("".join(str(int(v)) for v in row) for row in im) + "\n")
    print(out, im.shape, int(im.sum()))
