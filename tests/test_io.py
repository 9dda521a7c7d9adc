"""Save-point formats (lydion_amd/io.py) on the CPU: the static-cloud history keeps the
reference's dataset names, and its edge-cell index lists are those of SaveTimeStepHDF5
(src/SaveHDF5Data_StaticCloudTests.jl:176-200), in the order of the oracle's edge spectra."""
import numpy as np

from lydion_amd.io import StaticHistory
from oracle.oracle import OracleState, load
from tests.helpers import randomize, small_cloud


def test_edge_indices_follow_the_edge_spectra(tmp_path):
    cloud = randomize(small_cloud(n_r=9, n_z=7, n_nu=16), seed=4)
    o = OracleState(cloud)
    raw, norm = o.edge_spectra()
    ri, zi = StaticHistory.edge_indices(9, 7)
    assert len(ri) == raw.shape[0] == load().orc_edge_count(9, 7)
    J = cloud.lya["J"]
    for e in range(len(ri)):
        assert np.array_equal(raw[e], J[zi[e] - 1, ri[e] - 1])   # 1-based (i, j) of the reference
    h = StaticHistory(cloud.lya)
    for s in range(3):
        h.record(s * cloud.dt, o.step(), o.edge_spectra())
    path = tmp_path / "hist.npz"
    h.save(str(path))
    with np.load(path, allow_pickle=False) as f:
        for name in ("Times", "M_F_History", "M_F_radial_net_History", "M_F_radial_out_History", "L_Lya_History",
                     "f_esc_Lya_History", "J_Lya_EdgeSpectra", "J_Lya_EdgeSpectra_Norm", "J_Lya_Edge_rIndex",
                     "J_Lya_Edge_ZIndex", "ν", "νC", "δνC", "Δν"):
            assert name in f.files, name
        assert f["J_Lya_EdgeSpectra"].shape == (3, len(ri), 16)
        assert f["f_esc_Lya_History"].shape == (3,)
