"""Save points on the GPU: a run interrupted by a checkpoint (lydion_amd.io.save_checkpoint, the
J_Lya / H_r_Lya / H_Z_Lya of /Checkpoint) and resumed in a new context from it
(load_checkpoint, the restart of src/RHDSolver_SingleStar.jl:1359-1372) is bit-identical to the
uninterrupted run; e_Lyα is recomputed at the restart."""
import numpy as np
import pytest

import lydion_amd as L
from lydion_amd import m1lya as M
from lydion_amd.io import load_checkpoint, save_checkpoint
from lydion_amd.static_cloud import developed_state
from tests.helpers import bitdiff, small_cloud

pytestmark = pytest.mark.gpu


def test_checkpoint_restart_is_bit_identical(tmp_path):
    lut = L.VoigtLUT()
    cloud = developed_state(small_cloud(n_r=20, n_z=16, n_nu=64, tau0=1e7, tau_abs=0.5, umax=100.0, lut=lut))
    args = (cloud.c_tilde, cloud.dt, cloud.Zc)
    a = M.M1LyaContext(cloud.lya, cloud.geometry)
    M.run_steps(a, 5, *args, bc_z_bot=cloud.bc_z_bot)
    b = M.M1LyaContext(cloud.lya, cloud.geometry)
    M.run_steps(b, 3, *args, bc_z_bot=cloud.bc_z_bot)
    path = str(tmp_path / "ckpt.npz")
    save_checkpoint(path, b, cloud.c_tilde)
    b.close()
    c = M.M1LyaContext(cloud.lya, cloud.geometry)
    load_checkpoint(path, c)
    e = c.download("e_lya")
    J = c.download("J")
    ref_e = (4 * np.pi / L.constants.c) * np.sum(J * cloud.lya["dnu"], axis=-1)
    assert np.max(np.abs(e - ref_e) / np.maximum(ref_e, 1e-300)) <= 1e-12
    M.run_steps(c, 2, *args, bc_z_bot=cloud.bc_z_bot)
    for key in ("J", "H_r", "H_z"):
        assert bitdiff(c.download(key), a.download(key)) == 0, key
