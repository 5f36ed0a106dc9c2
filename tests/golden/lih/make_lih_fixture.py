"""Writes lih_matrices.npz from the reference demo's LiH inputs
(demo/LiH_example: DensityMatrix, SMatrix, DensityMatrixSecondary, text
matrices of 176 x 176) and copies AtomicCoords and the two basis files.
Run where the reference checkout is available; the fixtures are committed."""
import os
import shutil
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/demo/LiH_example"
HERE = os.path.dirname(os.path.abspath(__file__))

for f in ("AtomicCoords", "Li_gaussian_paulzim", "H_gaussian_paulzim"):
    shutil.copyfile(os.path.join(SRC, f), os.path.join(HERE, f))
np.savez_compressed(os.path.join(HERE, "lih_matrices.npz"),
                    density_matrix=np.loadtxt(os.path.join(SRC, "DensityMatrix")),
                    s_matrix=np.loadtxt(os.path.join(SRC, "SMatrix")),
                    density_matrix_secondary=np.loadtxt(os.path.join(SRC, "DensityMatrixSecondary")))
