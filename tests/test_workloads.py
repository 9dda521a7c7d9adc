"""The seeded sweep generator: reproducible, per-cloud seeds, draws in the stated ranges."""
import numpy as np

from lydion_amd import workloads as W
from lydion_amd.static_cloud import DEFAULT_SEED


def test_sweep_params_ranges_and_reproducibility():
    ps = W.sweep_params(64)
    assert ps == W.sweep_params(64)
    assert W.sweep_params(8) == ps[:8]            # cloud i does not depend on the sweep size
    assert [p["seed"] for p in ps] == [DEFAULT_SEED + i for i in range(64)]
    for p in ps:
        assert p["n_r"] in W.SIDES and p["n_z"] in W.SIDES and p["n_nu"] in W.BINS
        assert 1e5 <= p["tau0"] <= 1e9 and 1e2 <= p["T"] <= 1e4 and 0 <= p["tau_abs"] <= 10
        assert p["source"] in ("point", "uniform")
    # every choice occurs in 64 draws
    assert {p["n_r"] for p in ps} == set(W.SIDES) and {p["n_nu"] for p in ps} == set(W.BINS)
    assert {p["source"] for p in ps} == {"point", "uniform"}


def test_sweep_cloud_setup():
    ps = W.sweep_params(6)
    for p in ps:
        c = W.sweep_cloud(p)
        L = c.lya
        assert (L.n_r, L.n_z, L.n_nu) == (p["n_r"], p["n_z"], p["n_nu"])
        assert np.all(c.T == p["T"])
        js = L["js"]
        if p["source"] == "point":
            assert np.count_nonzero(js) == 1
        else:
            assert np.count_nonzero(js) > 1
        Lsrc = 4 * np.pi * (js * c.geometry.dV).sum() / c.params["L_lya"]
        if p["source"] == "point":
            assert abs(Lsrc - 1.0) < 1e-12              # the single-cell branch carries L
        else:
            # js = 3L/((4π)² R³) inside the sphere; the box holds its Z >= 0 half (reflecting bottom)
            assert abs(Lsrc - 0.5) < 0.05


def test_rhd_like_state_fields():
    c = W.rhd_like_cloud(32, 64, 32)
    L = c.lya
    assert c.T.min() == 100.0 and c.T.max() == 1e4                      # T spans 1e2..1e4 K
    nH = L["alpha"] / W_sigma0(c.T)
    assert np.allclose(nH.mean(), 1e3, rtol=1e-12)                        # unit-mean lognormal x 1e3
    assert np.log(nH).std() == pytest_approx(1.0, rel=1e-12)
    u = np.sqrt(np.mean(c.vel["u_r"] ** 2 + c.vel["u_z"] ** 2))
    assert u == pytest_approx(2.0 * c.params["c_s"], rel=1e-12)          # Mach 2
    assert np.all(c.vel["ur_star"][:, 0] == 0.0)                          # no flow through the axis
    dust = L["alpha_c"] / nH
    assert np.allclose(dust, dust.flat[0], rtol=1e-12)                    # α_c ∝ n_H
    assert np.allclose(L["alpha_F"], L["alpha"][..., None] * L["Hvoigt"] + L["alpha_c"][..., None], rtol=1e-14)
    assert np.count_nonzero(L["js"]) == 1 and c.bc_z_bot == 1


from pytest import approx as pytest_approx  # noqa: E402
from lydion_amd.lya_arrays import compute_sigma0 as W_sigma0  # noqa: E402
