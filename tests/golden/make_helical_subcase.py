"""Write tests/golden/helical_subcase.npz: the 64x64x32 sub-case of the
helical mixed-precision configuration (4 turns at pitch 1.0, magnification
2, voxel 0.5 ALU with slice aspect 2.5, detector 16 rows x 96 channels, 32
views per turn) -- its seeded ellipsoid phantom, the FP64 oracle's forward
projection plus 1% noise (float32), and the FP64 oracle's VCD reconstruction
after 6 iterations from np.random.seed(0) (default parameters: qGGMRF prior,
DC damping, auto-regularization).  CPU only; run from the repository root:
    python tests/golden/make_helical_subcase.py
"""

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import cone as oc  # noqa: E402
from oracle import parallel as op  # noqa: E402
from oracle import vcd  # noqa: E402
from mbirtorch_amd.phantom import random_ellipsoids, add_noise  # noqa: E402

TURNS, PER_TURN = 4, 32
V, R, C = TURNS * PER_TURN, 16, 96
SDD, SID = 368.0, 184.0
RECON_SHAPE = (64, 64, 32)
DV, SLICE_ASPECT = 0.5, 2.5
ITERS = 6


def geometry():
    angles = -np.pi + 2 * np.pi * np.arange(V) / PER_TURN
    h_iso = R * 1.0 / (SDD / SID)            # detector height at the isocentre
    travel = TURNS * h_iso                   # pitch 1.0
    z = np.linspace(-travel / 2, travel / 2, V)
    return angles, z


def main():
    t0 = time.time()
    angles, z = geometry()
    g = oc.geometry((V, R, C), angles, SDD, SID, recon_shape=RECON_SHAPE, delta_voxel=DV,
                    z_shifts=z, voxel_slice_aspect=SLICE_ASPECT)
    vol = random_ellipsoids(RECON_SHAPE, 6, seed=11, axes_range=(3.0, 14.0))
    idx = op.full_indices(RECON_SHAPE)
    clean = oc.forward(vol.reshape(-1, RECON_SHAPE[2])[idx], idx, g)
    sino = add_noise(clean, 0.01, seed=12).astype(np.float32)
    print(f'sinogram {sino.shape} in {time.time() - t0:.0f} s', flush=True)
    np.random.seed(0)
    ref, st = vcd.vcd_recon(vcd.ConeSystem(g), sino, {}, ITERS)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'helical_subcase.npz')
    np.savez_compressed(out, sino=sino, angles=angles, z_shifts=z, phantom=vol.astype(np.float32),
                        recon=ref.astype(np.float32), fm_rmse=np.asarray(st['fm_rmse']),
                        iters=ITERS, seed=0, sdd=SDD, sid=SID, delta_voxel=DV,
                        voxel_slice_aspect=SLICE_ASPECT)
    print(f'wrote {out} in {time.time() - t0:.0f} s; fm_rmse {st["fm_rmse"]}')


if __name__ == '__main__':
    main()
