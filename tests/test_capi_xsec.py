"""The library's own cross sections (pic_cross_section, host FP64) against the
oracle and the tabulated source: the same functions the collision kernels
evaluate, so a transcription error in the device tables fails here.  No GPU."""
import numpy as np
import pytest

import orc
import pic2d
from test_oracle_collisions import EXC_PINS, _fixture_excitation

AR, H = pic2d.GAS_ARGON, pic2d.GAS_HYDROGEN


@pytest.mark.parametrize("E,want", EXC_PINS)
def test_library_excitation_table_pins(E, want):
    assert pic2d.cross_section(AR, pic2d.XS_EXCITATION, E) == pytest.approx(want, abs=1e-9)


def test_library_excitation_matches_fixture_and_oracle():
    rng = np.random.default_rng(5)
    for E in np.concatenate([rng.uniform(0.0, 150.0, 500), [11.55, 11.624, 40.0, 120.0, 1e4]]):
        lib = pic2d.cross_section(AR, pic2d.XS_EXCITATION, E)
        assert lib == pytest.approx(_fixture_excitation(E) if E >= 11.55 else 0.0, abs=1e-9)
        assert lib == pytest.approx(orc.sigma("AR_EXC", E, 11.55)[0], abs=1e-9)


@pytest.mark.parametrize("ch,kind,thr", [
    (pic2d.XS_ELASTIC, "AR_EL_PHYS", 0.0), (pic2d.XS_IONIZATION, "AR_IZ_PHYS", 15.76),
    (pic2d.XS_CX, "AR_CX", 0.0), (pic2d.XS_ION_ISO, "AR_ISO", 0.0)])
def test_library_argon_fits_match_oracle(ch, kind, thr):
    E = np.geomspace(1e-3, 1e3, 200)
    lib = pic2d.cross_section(AR, ch, E)
    assert np.allclose(lib, orc.sigma(kind, E, thr), rtol=1e-12, atol=1e-12)


def test_reference_mode_fits_and_unsupported_channels():
    assert pic2d.cross_section(AR, pic2d.XS_ELASTIC, 0.1, pic2d.MODE_REF) == pytest.approx(-9.591, rel=2e-3)
    assert pic2d.cross_section(AR, pic2d.XS_IONIZATION, 16.0, pic2d.MODE_REF) == pytest.approx(0.158, rel=2e-3)
    with pytest.raises(pic2d.PicError) as e:
        pic2d.cross_section(AR, pic2d.XS_EXCITATION, 20.0, pic2d.MODE_REF)
    assert e.value.code == -7
    with pytest.raises(pic2d.PicError):
        pic2d.cross_section(H, pic2d.XS_EXCITATION, 20.0)
