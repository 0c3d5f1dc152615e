"""GPU FP64 multigrid (pic_mglin) against the oracle, bit for bit (the solver
keeps the reference's FP64 association).  The oracle itself is pinned bit for
bit against the reference's own Mglin on the CPU (test_oracle_mglin.py); no
reference object is loaded here."""
import numpy as np
import pytest

import orc
import pic2d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [5, 9, 17, 33, 65, 129, 257, 513])
@pytest.mark.parametrize("ncycle", [1, 2])
def test_mglin_bit_exact(n, ncycle):
    print(f"mglin n={n} ncycle={ncycle}", flush=True)
    rng = np.random.default_rng(7 * n + ncycle)
    f = rng.standard_normal((n, n)) * 1e5
    L = 0.0184
    g = pic2d.mglin(f, ncycle, L)
    o = orc.mglin(f, ncycle, L)
    assert np.array_equal(g, o), np.max(np.abs(g - o))


def test_mglin_bad_grid():
    with pytest.raises(pic2d.PicError) as e:
        pic2d.mglin(np.zeros((7, 7)))
    assert e.value.code == -2
