"""Pins of the oracle's Lyα source terms (PrepareLyαSubCycling!,
src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2086-2155):

  - Prob_Lyα_SH95 (src/LymanAlphaEmission.jl:79-91): the table values at its nodes, flat beyond
    the table, bilinear in (log T, log n_e) between nodes;
  - krecHII_A/B (src/PhotoChemistryAndCooling/GasChemistryRates.jl:166-204): the textbook
    hydrogen recombination coefficients at 1e4 K, α_A = 4.18e-13 and α_B = 2.59e-13 cm³ s⁻¹
    (Osterbrock & Ferland 2006, table 2.1), to the 1 % of the fits;
  - ∫dν_js_recLyα: P krec n_HII n_e E_Lyα/4π;  ∫dν_js_collLyα (:143-173): proportional to
    n_HI n_e, and at low T the n = 2 term's Boltzmann factor sets the temperature slope;
  - pd_Lyα (src/LymanAlphaDestruction.jl:43-97): in the collisional limit the 2s and 2p levels
    are mixed, so p_d → A_2γ / (3 A_Lyα + A_2γ); the H₂ term is capped at 1.3e-3 n_H2/n_HI;
  - the continuum opacities: additive in the dust species, dust below 1.1 ρd_floor ignored,
    Thomson scattering in the flux opacity only.
"""
import math

import numpy as np
import pytest

from lydion_amd import constants as K
from oracle import oracle as O

lib = O.load()


def test_sh95_nodes_clamp_and_bilinear():
    # nodes (src/LymanAlphaEmission.jl: row T = 1e4 K)
    assert lib.orc_prob_lya_sh95(1e4, 1e4, 1) == 0.92503048
    assert lib.orc_prob_lya_sh95(1e4, 1e4, 0) == 0.45373431
    assert lib.orc_prob_lya_sh95(5e2, 1e2, 1) == 0.80793637
    # flat beyond the table
    assert lib.orc_prob_lya_sh95(1e6, 1e4, 1) == lib.orc_prob_lya_sh95(3e4, 1e4, 1)
    assert lib.orc_prob_lya_sh95(1e4, 1e20, 1) == lib.orc_prob_lya_sh95(1e4, 1e14, 1)
    assert lib.orc_prob_lya_sh95(1e4, 1e12, 0) == lib.orc_prob_lya_sh95(1e4, 1e10, 0)
    assert lib.orc_prob_lya_sh95(10.0, 1.0, 1) == 0.80793637
    # the centre of a table cell (in log T, log n_e) is the mean of its corners
    T = 10 ** (0.5 * (math.log10(1e4) + math.log10(1.25e4)))
    ne = 10 ** 4.5
    mean = 0.25 * (0.92503048 + 0.9891268 + 0.9176273 + 0.98784252)
    assert lib.orc_prob_lya_sh95(T, ne, 1) == pytest.approx(mean, rel=1e-12)


def test_recombination_coefficients_at_1e4_K():
    assert lib.orc_krec_hii(1e4, 0) == pytest.approx(4.18e-13, rel=0.01)
    assert lib.orc_krec_hii(1e4, 1) == pytest.approx(2.59e-13, rel=0.01)
    assert lib.orc_krec_hii(1e4, 1) < lib.orc_krec_hii(1e4, 0)          # B excludes captures to n = 1
    assert lib.orc_krec_hii(0.1, 1) == lib.orc_krec_hii(1.0, 1)          # T clamped below 1 K


def test_recombination_emissivity():
    T, nHII, ne = 8e3, 3.0, 5.0
    for ots in (1, 0):
        ref = lib.orc_prob_lya_sh95(T, ne, ots) * lib.orc_krec_hii(T, ots) * nHII * ne * (
            K.h * K.nu_Lya / K.Four_pi)
        assert lib.orc_js_rec_lya(T, nHII, ne, ots) == pytest.approx(ref, rel=1e-15)


def test_collisional_emissivity_scaling():
    j = lib.orc_js_coll_lya(1.2e4, 2.0, 3.0)
    assert lib.orc_js_coll_lya(1.2e4, 4.0, 3.0) == pytest.approx(2 * j, rel=1e-15)
    assert lib.orc_js_coll_lya(1.2e4, 2.0, 6.0) == pytest.approx(2 * j, rel=1e-15)
    # at 2000 K the n = 2 term dominates: d ln(q √T) / d(1/T) -> -T_12 = -(3/4) E_Ryd / k_B
    T_12 = 0.75 * K.h * K.nu_LyC / K.kB
    T1, T2 = 2000.0, 2010.0
    q1 = lib.orc_js_coll_lya(T1, 1.0, 1.0) * math.sqrt(T1)
    q2 = lib.orc_js_coll_lya(T2, 1.0, 1.0) * math.sqrt(T2)
    slope = (math.log(q1) - math.log(q2)) / (1 / T1 - 1 / T2)
    assert slope == pytest.approx(-T_12, rel=2e-3)


def test_destruction_probability_limits():
    # collisional limit: 2s and 2p mixed, p_d -> A_2γ / (3 A_Lyα + A_2γ)
    big = 1e14
    pd = lib.orc_pd_lya(big, big, big, 0.0, 0.0, 1e4, 0.0)
    assert pd == pytest.approx(8.2206 / (3 * 6.265e8 + 8.2206), rel=1e-6)
    # grows with density below that limit
    p1 = lib.orc_pd_lya(1.0, 1.0, 1e3, 0.0, 0.0, 1e4, 0.0)
    p2 = lib.orc_pd_lya(1e3, 1e3, 1e6, 0.0, 0.0, 1e4, 0.0)
    assert p1 < p2 < pd
    # the H2 term: proportional to n_H2/n_HI, capped at 1.3e-3 n_H2/n_HI (hot gas); none below n_HI = 1e-20
    base = lib.orc_pd_lya(1.0, 1.0, 1e4, 0.0, 0.0, 1e5, 0.0)
    with_h2 = lib.orc_pd_lya(1.0, 1.0, 1e4, 0.0, 1.0, 1e5, 0.0)   # n_H2/n_HI = 1e-4, hot: at the cap
    assert with_h2 - base == pytest.approx(1.3e-3 * 1e-4, rel=1e-6)
    cold = lib.orc_pd_lya(1.0, 1.0, 1e4, 0.0, 1.0, 300.0, 0.0) - lib.orc_pd_lya(1.0, 1.0, 1e4, 0.0, 0.0, 300.0, 0.0)
    assert 0.0 < cold < 1e-6 * 1.3e-3 * 1e-4                        # exp(-14540/300): far below the cap
    no_hi = lib.orc_pd_lya(1.0, 1.0, 1e-21, 0.0, 1.0, 1e5, 0.0) - lib.orc_pd_lya(1.0, 1.0, 1e-21, 0.0, 0.0, 1e5, 0.0)
    assert abs(no_hi) < 1e-12                                       # n_HI <= 1e-20: no H2 term
    assert lib.orc_pd_lya(1e30, 1e30, 1e30, 1e30, 1e30, 1e4, 0.0) <= 1.0


def test_source_opacities():
    n = 4
    chem = np.zeros((8, 1, n))
    chem[0] = 1e4
    chem[1] = 1e2          # n_HI
    chem[2] = 3.0          # n_H2
    chem[5] = 2.0          # n_e
    chem[6] = 0.5          # n_OHx
    chem[7] = 0.25         # n_CHx
    rho_d = np.zeros((1, n, 3))
    rho_d[0, 1] = [1e-20, 0.0, 0.0]
    rho_d[0, 2] = [0.0, 2e-20, 3e-20]
    rho_d[0, 3] = [1e-40, 0.0, 0.0]      # below 1.1 ρd_floor: ignored
    kap = np.array([1e4, 2e4, 3e4, 5e4, 6e4, 7e4])
    out = O.lya_sources(chem, rho_d, kap, destruction=False)
    gas = 3.0 * 2.1e-25 + 0.5 * 4.07e-18 + 0.25 * 0.05e-18
    assert out[1, 0, 0] == pytest.approx(gas, rel=1e-15)
    assert out[2, 0, 0] == pytest.approx(gas + 2.0 * 6.6524587321e-25, rel=1e-15)
    assert out[1, 0, 1] - out[1, 0, 0] == pytest.approx(1e4 * 1e-20, rel=1e-9)
    assert out[2, 0, 2] - out[2, 0, 0] == pytest.approx(6e4 * 2e-20 + 7e4 * 3e-20, rel=1e-9)
    assert out[1, 0, 3] == out[1, 0, 0]
    assert np.all(out[3] == 0.0)                                   # no destruction requested
    js = out[0, 0, 0]
    assert js == pytest.approx(lib.orc_js_rec_lya(1e4, 0.0, 2.0, 1) + lib.orc_js_coll_lya(1e4, 1e2, 2.0), rel=1e-15)
