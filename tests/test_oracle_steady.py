"""Reference steady-state walls in the oracle (orc32_steady), pinned against a
literal restatement of the reference's scan (boundary.cpp:890-1082): for every
electron outside a side, in array order, search the ion array from a cursor
for an ion outside the same side; on a hit both are re-injected and the cursor
moves past the ion.  Run in id order on both arrays, that scan must pick
exactly the pairs orc32_steady re-injects (same new position for the two
members of a pair), x sides before y sides."""
import numpy as np

import orc

N = 17


def sides_x(jx):
    return 0 if jx < 0 else (1 if jx >= N else -1)


def sides_y(jy):
    return 2 if jy < 0 else (3 if jy >= N else -1)


def reference_scan(e, i):
    """the reference's cursor loops on id-ordered arrays; returns pairs (eid, iid)"""
    pairs = []
    moved_e, moved_i = set(), set()
    for side_of, sides in ((lambda c: sides_x(c[0]), (0, 1)), (lambda c: sides_y(c[1]), (2, 3))):
        ce = [k for k in range(len(e["id"])) if k not in moved_e]
        ci = [k for k in range(len(i["id"])) if k not in moved_i]
        ce.sort(key=lambda k: e["id"][k])
        ci.sort(key=lambda k: i["id"][k])
        cur = {s: 0 for s in sides}
        amount = {s: sum(1 for k in ce if side_of(e["c"][k]) == s) for s in sides}
        for k in ce:
            s = side_of(e["c"][k])
            if s not in sides:
                continue
            j = cur[s]
            while amount[s] > 0 and j < len(ci):
                if side_of(i["c"][ci[j]]) == s:
                    pairs.append((int(e["id"][k]), int(i["id"][ci[j]])))
                    moved_e.add(k)
                    moved_i.add(ci[j])
                    amount[s] -= 1
                    break
                j += 1
            cur[s] = j + 1
    return pairs


def make(n, rng, id0):
    jx = rng.integers(-3, N + 3, n)
    jy = rng.integers(-3, N + 3, n)
    ids = rng.permutation(np.arange(id0, id0 + 3 * n, dtype=np.uint32))[:n]
    st = dict(cell=orc.cell_pack(jx, jy), ox=rng.random(n).astype(np.float32),
              oy=rng.random(n).astype(np.float32),
              vx=rng.standard_normal(n).astype(np.float32),
              vy=rng.standard_normal(n).astype(np.float32),
              vz=rng.standard_normal(n).astype(np.float32), id=ids)
    return st, dict(id=ids, c=list(zip(jx.tolist(), jy.tolist())))


def run_steady(e, i, step=3, key=(11, 22), se=2.0, si=0.5):
    e = {k: v.copy() for k, v in e.items()}
    i = {k: v.copy() for k, v in i.items()}
    P = orc.P
    m = orc.lib().orc32_steady(len(e["id"]), P(e["cell"]), P(e["ox"]), P(e["oy"]), P(e["vx"]),
                               P(e["vy"]), P(e["vz"]), P(e["id"]), len(i["id"]), P(i["cell"]),
                               P(i["ox"]), P(i["oy"]), P(i["vx"]), P(i["vy"]), P(i["vz"]),
                               P(i["id"]), N, step, key[0], key[1], se, si)
    return m, e, i


def test_pairing_matches_the_reference_scan():
    rng = np.random.default_rng(7)
    for trial in range(20):
        e, eref = make(int(rng.integers(5, 60)), rng, 0)
        i, iref = make(int(rng.integers(5, 60)), rng, 1000)
        m, e1, i1 = run_steady(e, i)
        pairs = reference_scan(eref, iref)
        assert m == len(pairs)
        emoved = {int(e["id"][k]) for k in range(len(e["id"])) if e1["cell"][k] != e["cell"][k] or
                  e1["ox"][k] != e["ox"][k]}
        imoved = {int(i["id"][k]) for k in range(len(i["id"])) if i1["cell"][k] != i["cell"][k] or
                  i1["ox"][k] != i["ox"][k]}
        assert emoved == {a for a, _ in pairs} and imoved == {b for _, b in pairs}
        epos = {int(e["id"][k]): (e1["cell"][k], e1["ox"][k], e1["oy"][k]) for k in range(len(e["id"]))}
        ipos = {int(i["id"][k]): (i1["cell"][k], i1["ox"][k], i1["oy"][k]) for k in range(len(i["id"]))}
        for a, b in pairs:   # a pair lands on one point inside [dx, L - dx]^2
            assert epos[a] == ipos[b]
            jx, jy = orc.cell_unpack(np.array([epos[a][0]], dtype=np.uint32))
            assert 1 <= jx[0] <= N - 3 and 1 <= jy[0] <= N - 3


def test_reinjected_velocities_follow_the_reference_sampler():
    # v = vt cos(2 pi R2) sqrt|ln R1|: mean 0, variance vt^2 / 2
    rng = np.random.default_rng(8)
    n = 20000
    e = dict(cell=orc.cell_pack(np.full(n, -1), np.full(n, 5)), ox=np.full(n, 0.5, np.float32),
             oy=np.full(n, 0.5, np.float32), vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32),
             vz=np.zeros(n, np.float32), id=np.arange(n, dtype=np.uint32))
    i = {k: v.copy() for k, v in e.items()}
    i["id"] = np.arange(n, 2 * n, dtype=np.uint32)
    m, e1, i1 = run_steady(e, i, se=3.0, si=0.25)
    assert m == n
    for a, vt in ((e1, 3.0), (i1, 0.25)):
        for k in ("vx", "vy", "vz"):
            v = a[k].astype(np.float64)
            assert abs(v.mean()) < 5 * vt / np.sqrt(2 * n)
            assert abs(v.var() / (vt * vt / 2) - 1) < 0.05
