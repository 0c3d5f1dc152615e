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


# ------------------------------------------------- e-Ar excitation (physical)
# Hand sums of the six excitation tables (2)-(7) of the recommended Ar set in
# the reference's References/PhelpsAr_e_cross_section.pdf at energies where
# every table has a row (or is 0 / held), in 1e-20 m^2:
#   15 eV: .022 + .028 + .0045 + .065 + .15 + .028    = 0.2975
#   20 eV: .033 + .047 + .0066 + .103 + .50 + .105    = 0.7946
#   50 eV: .0036 + .045 + .0012 + .175 + .43 + .235   = 0.8898
#  100 eV: .0006 + .031 + .0005 + .15 + .28 + .27     = 0.7321
#  200 eV (every table held at its last row):
#          .0005 + .031 + .0004 + .15 + .25 + .26     = 0.6919
#  13 eV: .009 + .012 + .0018 + .029 + .003 + 0       = 0.0548
EXC_PINS = [(15.0, 0.2975), (20.0, 0.7946), (50.0, 0.8898), (100.0, 0.7321), (200.0, 0.6919),
            (13.0, 0.0548)]


def _fixture_excitation(E):
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "ar_excitation_tables.json")))
    tot = 0.0
    for t in d["excitation"]:
        e = np.array([r[0] for r in t["rows"]])
        s = np.array([r[1] for r in t["rows"]])
        if E >= e[0]:
            tot += s[-1] if E >= e[-1] else float(np.interp(E, e, s))
    return tot


@pytest.mark.parametrize("E,want", EXC_PINS)
def test_excitation_pinned_to_tabulated_source(E, want):
    assert orc.sigma("AR_EXC", E, 11.55)[0] == pytest.approx(want, abs=1e-12)
    assert _fixture_excitation(E) == pytest.approx(want, abs=1e-12)


def test_excitation_threshold_and_fixture_agreement():
    # nothing below the lumped channel's 11.55 eV threshold
    for E in (0.0, 5.0, 11.5, 11.549, 11.5499):
        assert orc.sigma("AR_EXC", E, 11.55)[0] == 0.0
    # between rows: linear in every table, so the oracle's per-table sum equals
    # the fixture's interpolation everywhere
    rng = np.random.default_rng(3)
    for E in np.concatenate([rng.uniform(11.55, 130.0, 400), [11.55, 12.25, 14.09, 120.0]]):
        assert orc.sigma("AR_EXC", E, 11.55)[0] == pytest.approx(_fixture_excitation(E),
                                                                 rel=1e-12, abs=1e-15)
    # the peak of the lumped channel: 0.9172 at 40 eV (tables' union grid)
    E = np.linspace(11.55, 150.0, 3000)
    s = orc.sigma("AR_EXC", E, 11.55)
    assert 35.0 < E[np.argmax(s)] < 45.0 and s.max() == pytest.approx(0.9172, abs=2e-4)
