"""Save points of the device-resident state, in the reference's dataset names.

The reference writes HDF5 through HDF5.jl: the static-cloud histories and edge spectra
(src/SaveHDF5Data_StaticCloudTests.jl:17-425) and the RHD restart group /Checkpoint with
J_Lya, H_r_Lya, H_Z_Lya (src/SaveHDF5Data_RHD_SingleStar.jl:157-200,251-261,1317-1319), read
back at a restart (src/RHDSolver_SingleStar.jl:1359-1372).  Julia column-major [N_ν, N_r, N_Z]
arrays appear in HDF5 (row-major) as (N_Z, N_r, N_ν), which is the layout of the arrays here, so
the names and shapes below are what a reader of the reference's files sees.  Neither h5py nor an
HDF5 library is available to this package, so the files are NumPy .npz archives with those names
("Checkpoint/J_Lya", ...); the Julia binding keeps the reference's own HDF5 writers and calls
download_state! / upload_state! around them instead.
"""
from __future__ import annotations

import numpy as np

from . import constants as K

CHECKPOINT_FIELDS = (("J_Lya", "J"), ("H_r_Lya", "H_r"), ("H_Z_Lya", "H_z"))


def save_checkpoint(path: str, ctx, c_tilde: float | None = None) -> None:
    """The Lyα part of /Checkpoint: J, H_r, H_Z of the device state (and the frequency grid and
    c̃ that GenerateLyαArrays_FromSaved needs at a restart)."""
    lya = ctx.lya
    data = {f"Checkpoint/{name}": ctx.download(key) for name, key in CHECKPOINT_FIELDS}
    data["ν"] = np.asarray(lya.nu, dtype=np.float64)
    data["νC"] = np.asarray(lya["nuC"], dtype=np.float64)
    data["Δν"] = np.asarray(lya["dnu"], dtype=np.float64)
    if c_tilde is not None:
        data["Checkpoint/c_tilde_Lya"] = np.float64(c_tilde)
    np.savez(path, **data)


def load_checkpoint(path: str, ctx) -> None:
    """The restart of src/RHDSolver_SingleStar.jl:1359-1372: J and H from /Checkpoint, the *_old
    copies equal to them, and e_Lyα recomputed (Update_e_Lyα!), all on the device."""
    from . import m1lya as M
    with np.load(path, allow_pickle=False) as f:
        arrs = {key: f[f"Checkpoint/{name}"] for name, key in CHECKPOINT_FIELDS}
    for key, a in arrs.items():
        ctx.upload(key, a)
        ctx.upload(key + "_old", a)
    M.update_e_lya(ctx)


class StaticHistory:
    """The time series SaveTimeStepHDF5 keeps for an RT-only static cloud
    (src/SaveHDF5Data_StaticCloudTests.jl:122-175): per saved time the force multipliers, L_Lyα,
    f_esc and the edge spectra (raw and normalised) with their cell indices."""

    def __init__(self, lya):
        self.lya = lya
        self.times: list[float] = []
        self.scalars: list[np.ndarray] = []
        self.edge_raw: list[np.ndarray] = []
        self.edge_norm: list[np.ndarray] = []

    def record(self, t: float, step_scalars, edge=None) -> None:
        """step_scalars: the 8 values of m1lya_step's out; edge: (raw, norm) edge spectra."""
        self.times.append(float(t))
        self.scalars.append(np.asarray(step_scalars, dtype=np.float64)[:5].copy())
        if edge is not None:
            self.edge_raw.append(np.asarray(edge[0], dtype=np.float32))
            self.edge_norm.append(np.asarray(edge[1], dtype=np.float32))

    @staticmethod
    def edge_indices(n_r: int, n_z: int):
        """1-based (i, j) of the edge cells in the order of the edge spectra, as SaveTimeStepHDF5
        lists them (src/SaveHDF5Data_StaticCloudTests.jl:176-200): the outer column i = N_r - 1 for
        j = 2 .. N_Z - 1, the top row j = N_Z - 1 for i = 2 .. N_r - 2, the bottom row j = 2."""
        ri, zi = [], []
        for j in range(2, n_z):            # outer column i = N_r - 1, j = 2 .. N_Z - 1
            ri.append(n_r - 1)
            zi.append(j)
        for i in range(2, n_r - 1):        # top row j = N_Z - 1, i = 2 .. N_r - 2
            ri.append(i)
            zi.append(n_z - 1)
        for i in range(2, n_r - 1):        # bottom row j = 2
            ri.append(i)
            zi.append(2)
        return np.array(ri, dtype=np.int32), np.array(zi, dtype=np.int32)

    def save(self, path: str) -> None:
        lya = self.lya
        s = np.array(self.scalars) if self.scalars else np.zeros((0, 5))
        dnuD = float(np.asarray(lya["dnuD"]).reshape(-1)[0]) if "dnuD" in lya.arrays else 1.0
        data = {"Times": np.array(self.times), "M_F_History": s[:, 0], "M_F_radial_net_History": s[:, 1],
                "M_F_radial_out_History": s[:, 2], "L_Lya_History": s[:, 3], "f_esc_Lya_History": s[:, 4],
                "ν": np.asarray(lya.nu, dtype=np.float64), "νC": np.asarray(lya["nuC"], dtype=np.float64),
                "δνC": (np.asarray(lya["nuC"]) - K.nu_Lya) / dnuD, "Δν": np.asarray(lya["dnu"], dtype=np.float64)}
        if self.edge_raw:
            data["J_Lya_EdgeSpectra"] = np.stack(self.edge_raw)
            data["J_Lya_EdgeSpectra_Norm"] = np.stack(self.edge_norm)
            data["J_Lya_Edge_rIndex"], data["J_Lya_Edge_ZIndex"] = self.edge_indices(lya.n_r, lya.n_z)
        np.savez(path, **data)
