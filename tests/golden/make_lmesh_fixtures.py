"""Copies the reference's unstructured-mesh data files into tests/golden/.

  sedov15oct.lmesh        as it is (875 elements, 1126 nodes)
  sedov21oct.lmesh.gz     gzip of sedov21oct.lmesh (31408 elements)
  sedov21oct_xyz.npz      xyz.asc -- the node coordinates at t = 1e-2 that
                          the reference's CUDA port writes for sedov21oct
                          (reference/CUDA/lulesh-kepler-singlegpu/lulesh.cu:
                          3449-3466) -- as float64 [num_node, 3]

Run where the reference tree is present:
    python tests/golden/make_lmesh_fixtures.py
"""
import gzip
import os
import shutil

import numpy as np

SRC = "/root/reference/reference/CUDA/lulesh-kepler-singlegpu"
DST = os.path.dirname(os.path.abspath(__file__))


def main():
    shutil.copyfile(os.path.join(SRC, "sedov15oct.lmesh"),
                    os.path.join(DST, "sedov15oct.lmesh"))
    with open(os.path.join(SRC, "sedov21oct.lmesh"), "rb") as f, \
            gzip.open(os.path.join(DST, "sedov21oct.lmesh.gz"), "wb", 9) as g:
        shutil.copyfileobj(f, g)
    vals = np.loadtxt(os.path.join(SRC, "xyz.asc"))
    rec = vals.reshape(-1, 4)
    assert np.array_equal(rec[:, 0], np.arange(len(rec)))
    np.savez_compressed(os.path.join(DST, "sedov21oct_xyz.npz"), xyz=rec[:, 1:].copy())


if __name__ == "__main__":
    main()
