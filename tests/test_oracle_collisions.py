"""Oracle cross sections against values computed from the reference's fits
(e_collisions.cpp:205-256, ion_collisions.cpp:170-176) and their documented
quirks: negative elastic values below 0.35 eV, undefined breakpoints."""
import numpy as np
import pytest

import orc


@pytest.mark.parametrize("kind,E,want", [
    ("AR_EL_REF", 1.0, 1.37), ("AR_EL_REF", 10.0, 20.945), ("AR_EL_REF", 50.0, 8.487),
    ("AR_IZ_REF", 16.0, 0.158), ("AR_IZ_REF", 50.0, 4.167), ("AR_CX", 1.0, 77.0),
    ("AR_CX", 10.0, 37.39), ("AR_EL_REF", 0.1, -9.591), ("AR_EL_REF", 0.34, -858.2),
])
def test_reference_fit_values(kind, E, want):
    assert orc.sigma(kind, E)[0] == pytest.approx(want, rel=2e-3)


def test_reference_breakpoints_and_physical_clamps():
    # exactly on a breakpoint the reference returns an uninitialised value: 0 here
    assert orc.sigma("AR_EL_REF", 0.35)[0] == 0.0 and orc.sigma("AR_EL_REF", 15.0)[0] == 0.0
    assert orc.sigma("AR_IZ_REF", 15.755)[0] == 0.0 and orc.sigma("AR_IZ_REF", 100.0)[0] == 0.0
    # the physical mode closes the branches and clamps the negative dip
    E = np.linspace(0.0, 0.35, 200)
    assert np.all(orc.sigma("AR_EL_PHYS", E) >= 0)
    assert orc.sigma("AR_EL_PHYS", 15.0)[0] == pytest.approx(126.2 * 15 ** -0.69)
    # the piecewise fits are continuous where they should be
    assert orc.sigma("AR_CX", 4.0 + 1e-9)[0] == pytest.approx(orc.sigma("AR_CX", 4.0)[0], rel=1e-6)
    assert orc.sigma("AR_IZ_PHYS", 15.76, 15.76)[0] >= 0.0


def test_phelps_isotropic_and_janev_forms():
    # Phelps (1994): 2e-19/(sqrt(e)(1+e)) + 3e-19 e/(1+e/3)^2 m^2 at 1 eV = 26.875e-20 m^2
    assert orc.sigma("AR_ISO", 1.0)[0] == pytest.approx(26.875)
    # Janev H ionisation vanishes below 13.6 eV and peaks near 50-60 eV
    E = np.linspace(14, 300, 300)
    s = orc.sigma("H_IZ", E)
    assert orc.sigma("H_IZ", 13.0)[0] == 0.0 and 40 < E[np.argmax(s)] < 80


def test_null_collision_bound_for_the_argon_workload():
    me, mi = 9.10939e-31, 39.948 * 1.66054e-27
    ng = 13.332 / (1.38066e-23 * 300.0)           # 100 mTorr at 300 K
    dt = 0.1 / np.sqrt(1e15 * 2.566971e-38 / (8.854e-12 * me))
    nue = orc.nu_max(me, ng, True)
    nui = orc.nu_max(mi, ng, False)
    assert 0.01 < nue * dt < 0.5                  # a tenth of the electrons are candidates
    assert nui * dt < 0.01
