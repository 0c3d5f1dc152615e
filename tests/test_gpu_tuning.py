"""The launch-shape choices do not change results: the same run with every
species sorted on every pass and with the tuned intervals (set on the live
run through pic_set_tuning) ends with the same particles (by id), charge,
phi and E, bit for bit; the tolerance solve meets its target with 1, 2 or 3
sweeps per leg."""
import numpy as np
import pytest

import pic2d
import workloads

pytestmark = pytest.mark.gpu


def final(p, tuning):
    with pic2d.Pic(p) as sim:
        sim.run(1)
        if tuning:
            sim.set_tuning(*tuning)
        sim.advance(11)
        out = {}
        for s in (0, 1):
            st = sim.get_state(s)
            o = np.argsort(st["id"], kind="stable")
            out.update({f"{s}.{k}": v[o] for k, v in st.items()})
            out[f"{s}.q"] = sim.get_qfixed(s)
        out["phi"] = sim.get_phi()
        out["E32"] = sim.get_E32()
        return out, sim.tuning()


@pytest.mark.parametrize("collisions", [True, False])
def test_sort_intervals_do_not_change_results(collisions):
    """collisionless, the species' in-place passes of a step share one launch
    (k_push_inplace_pair), with the next sort's histogram when both sort on
    the following step (intervals 3, 3), and run one by one when only one of
    them is in place (5, 3)"""
    p = workloads.discharge(64, 40, te=4.0)
    if collisions:
        p = workloads.with_argon(p, pressure_pa=60.0)
    p.mg_mode = pic2d.MG_FIXED
    a, ta = final(p, None)
    runs = [final(p, t) for t in ((1, 1, 0), (5, 3, 0), (3, 3, 0))]
    assert runs[0][1]["sort_interval_e"] == 1 and runs[1][1]["sort_interval_i"] == 3
    print("tunings", ta, [t for _, t in runs], flush=True)
    for b, _ in runs:
        for key in a:
            assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("nu", [1, 2, 3])
def test_sweeps_meet_the_target(nu):
    p = workloads.discharge(128, 10, te=4.0)
    p.mg_tol = 1e-10
    with pic2d.Pic(p) as sim:
        sim.run(1)
        sim.set_tuning(0, 0, nu)
        assert sim.tuning()["mg_sweeps"] == nu
        sim.advance(3)
        _, _, rr = sim.mg_stats()
        assert rr <= 1e-10
    with pytest.raises(pic2d.PicError):
        with pic2d.Pic(p) as sim:
            sim.set_tuning(0, 0, 4)
