"""Physics-level pins of the CPU oracle: static spheres run for 4 diffusion times, as the
reference's own validation runs them (RunLyaEscapeFractionStaticCloudTest.jl: T = 1e4 K, no
recoil, p_d = 0, reflecting bottom), compared with the tables and closed forms the reference
validates against:

  - f_esc of a dusty sphere against the Neufeld analytic solution and the Monte Carlo table
    (EscapeFractionTest_StaticCloud/Analytical*.txt, f_esc_tdMAX1.0_*_MCRT.txt), uniform and
    point source;
  - M_F / (aτ)^{1/3} of the dusty uniform source against AnalyticalMF_UniformSource.txt;
  - the emergent spectrum of the dust-free point source against J_sphere_point
    (src/ConstantsAndFunctions.jl:1018-1025), with f_esc → 1.

The grid is 32×32 with 64 bins at τ0 = 1e6 (aτ0 = 472), coarse against the reference's
128×128 at τ0 = 1e8, so the agreement is at the few-percent level; each tolerance is written
beside its check, and each is tighter than what a factor-2 slip in the dust opacity, the
Voigt parameter or the source luminosity would give (those move f_esc by ≥ 10 %, the
spectrum's mean |x| by ≥ 26 %, M_F by ≥ 12 %).
"""
import math

import numpy as np
import pytest

from lydion_amd.lya_arrays import compute_a
from oracle.oracle import OracleState
from tests import physics_cases as P

TAU0, N, N_NU = 1e6, 32, 64
ATAU13 = (compute_a(1e4) * TAU0) ** (1.0 / 3.0)


def _run(x_dust, source):
    cloud = P.sphere(TAU0, N, N_NU, x_dust, source)
    o = OracleState(cloud)
    hist = np.array([o.step() for _ in range(P.n_steps(cloud))])
    return cloud, o, hist


@pytest.fixture(scope="module")
def uniform_dusty():
    return _run(0.3, "uniform")


@pytest.fixture(scope="module")
def point_dusty():
    return _run(0.3, "point")


@pytest.fixture(scope="module")
def point_dust_free():
    return _run(0.0, "point")


def _final_finite(col):
    good = np.isfinite(col)
    return float(col[good][-1])


def test_uniform_source_escape_fraction(uniform_dusty):
    _, _, hist = uniform_dusty
    f = _final_finite(hist[:, 4])
    ana, mc = P.analytic_f_esc("uniform", 0.3), P.mcrt_f_esc("uniform", 0.3)
    print(f"uniform, (aτ)^1/3 τ_abs = 0.3: f_esc {f:.4f}, analytic {ana:.4f} ({f / ana - 1:+.2%}), "
          f"MCRT {mc:.4f} ({f / mc - 1:+.2%})")
    assert abs(f / ana - 1) < 0.06   # 32x32, aτ0 = 472: -3.7 %
    assert abs(f / mc - 1) < 0.08


def test_uniform_source_force_multiplier(uniform_dusty):
    _, _, hist = uniform_dusty
    mf = _final_finite(hist[:, 0]) / ATAU13
    ana = P.analytic_mf_uniform(0.3)
    print(f"uniform: M_F/(aτ)^1/3 {mf:.4f}, analytic {ana:.4f} ({mf / ana - 1:+.2%})")
    assert abs(mf / ana - 1) < 0.06  # +3.6 %


def test_point_source_escape_fraction(point_dusty):
    _, _, hist = point_dusty
    f = _final_finite(hist[:, 4])
    ana, mc = P.analytic_f_esc("point", 0.3), P.mcrt_f_esc("point", 0.3)
    print(f"point, (aτ)^1/3 τ_abs = 0.3: f_esc {f:.4f}, analytic {ana:.4f} ({f / ana - 1:+.2%}), "
          f"MCRT {mc:.4f} ({f / mc - 1:+.2%})")
    assert abs(f / ana - 1) < 0.08
    assert abs(f / mc - 1) < 0.08


def test_point_source_emergent_spectrum(point_dust_free):
    cloud, o, hist = point_dust_free
    f = _final_finite(hist[:, 4])
    assert 0.98 < f <= 1.0 + 1e-9  # dust-free: every photon escapes once the field is steady
    x, dx, S = P.escaping_spectrum(cloud, o["H_r"], o["H_z"])
    A = P.j_sphere_point(x, ATAU13 ** 3)
    A = A / np.sum(A * dx)
    l1 = float(np.sum(np.abs(S - A) * dx))
    xm, xa = float(np.sum(np.abs(x) * S * dx)), float(np.sum(np.abs(x) * A * dx))
    pk, pa = x[np.argmax(S * (x > 0))], x[np.argmax(A * (x > 0))]
    print(f"point, dust-free: f_esc {f:.5f}; spectrum L1 {l1:.3f}, mean |x| {xm:.3f} vs {xa:.3f} "
          f"({xm / xa - 1:+.2%}), red peak at x = {pk:.3f} vs {pa:.3f} (bin centres)")
    assert l1 < 0.2            # 0.137
    assert abs(xm / xa - 1) < 0.1  # +6.6 %
    assert pk == pa
    # symmetric about the line centre without recoil
    assert abs(float(np.sum(x * S * dx))) < 1e-6 * xm
