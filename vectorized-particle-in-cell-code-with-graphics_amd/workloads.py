"""Named workloads: the discharge configurations this library is benchmarked on.

Each function returns a :class:`pic2d.Params`.  Physical constants are the
reference's (constants.h:15-36); the grid is given in nodes (2^k + 1).

* ``plain``      -- 64x64 cells (65^2 nodes), 1 cm, 100 particles/cell/species,
                    Te = 2 eV, Ti = 0.026 eV, n = 1e15 m^-3, collisionless,
                    absorbing walls, w_pe dt = 0.1.
* ``argon_256``  -- 256x256 cells (257^2 nodes), 200 particles/cell/species
                    (13.1 M each), the same initial state, in argon at
                    100 mTorr / 300 K: null-collision MCC (e-Ar elastic,
                    11.55 eV excitation, 15.76 eV ionisation; Ar+-Ar isotropic
                    elastic and charge exchange) and an electron-ion pair
                    re-injected for every ion lost to a wall.
* ``large_1024`` -- 1024x1024 cells (1025^2 nodes), 100 particles/cell/species
                    (104.9 M each), collisionless.
* ``argon_1024`` -- the 1024x1024 / 100 ppc grid with the argon background.
"""
import math

import pic2d

CHARGE = 1.602177e-19
ELECTRON_MASS = 9.10939e-31
AMU = 1.66054e-27
EPSILON = 8.854e-12
ELECTRON_CHARGE2 = 2.566971e-38
AR_MASS = 39.948 * AMU
H_MASS = 1.00728 * AMU
SEED = 0x5EED0001
BOLTZMANN = 1.38066e-23


def discharge(cells, ppc, length=0.01, te=2.0, ti=0.026, density=1e15, ion_mass=AR_MASS,
              mode=pic2d.MODE_PHYSICAL, mg_mode=pic2d.MG_TOL, seed=SEED):
    p = pic2d.default_params()
    N = cells + 1
    n = cells * cells * ppc
    p.nodes = N
    p.init_particle_no = n
    p.mode = mode
    p.loader = pic2d.LOAD_RANDOM
    p.wall = pic2d.WALL_ABSORB
    p.have_collisions = 0
    p.length = length
    wpe = math.sqrt(density * ELECTRON_CHARGE2 / (EPSILON * ELECTRON_MASS))
    p.dt = 0.1 / wpe
    p.dti = p.dt
    p.dt_dti = 1
    p.ion_mass = ion_mass
    w = density * length * length / n          # real particles per macro-particle (1 m depth)
    p.qce = -CHARGE * w
    p.qci = CHARGE * w
    p.vte = math.sqrt(te * CHARGE / ELECTRON_MASS)
    p.vti = math.sqrt(ti * CHARGE / ion_mass)
    p.vtn = math.sqrt(0.025 * CHARGE / ion_mass)
    p.sigma_load_e = p.vte                     # 3-V Maxwellian, sigma = sqrt(kT/m)
    p.sigma_load_i = p.vti
    p.mg_mode = mg_mode
    p.mg_ncycle = 2
    p.mg_tol = 1e-8
    p.seed = seed
    p.run_time = 1000
    p.graph_steps = 0        # batch diagnostics off unless asked for (with_diagnostics)
    p.time_data_freq = 0
    p.reduce = 0             # fixed particle budget: no population reduction
    return p


def with_diagnostics(p, start, steps, time_data_freq=100, no_bins=30001, bins_per_vt=300):
    """Batch-mode graph averages over [start, start + steps) and timeData rows,
    at the reference's locations for this grid (parameters.h:288-313)."""
    N = p.nodes
    p.graph_start = start
    p.graph_steps = steps
    p.time_data_freq = time_data_freq
    p.no_bins = no_bins
    p.divide_bins_by = bins_per_vt
    p.dist_loc[0] = int(0.9 * N)
    p.dist_loc[1] = int(0.5 * N)
    p.dist_loc[2] = int(0.1 * N)
    p.dist_loc[3] = int(0.5 * N)
    p.spread = int(0.05 * N)
    p.probe[0], p.probe[1] = 3, N // 2
    p.probe[2], p.probe[3] = N // 2, N // 2
    p.probe[4], p.probe[5] = N - 3, N // 2
    p.write_graphs = 1
    return p


def plain(**kw):
    """64x64 cells, 100 ppc, collisionless (409 600 particles per species)."""
    return discharge(64, 100, **kw)


def with_argon(p, pressure_pa=13.332, temperature=300.0):
    """Argon background (default 100 mTorr, 300 K), physical-mode collisions
    and steady-state pair re-injection on ion loss."""
    p.have_collisions = 1
    p.gas = pic2d.GAS_ARGON
    p.neutral_density = pressure_pa / (BOLTZMANN * temperature)
    p.vtn = math.sqrt(BOLTZMANN * temperature / p.ion_mass)
    p.wall = pic2d.WALL_STEADY
    return p


def with_gas(p, gas, pressure_pa, temperature=300.0):
    """Argon or hydrogen background, physical-mode collisions and steady-state
    pair re-injection (hydrogen: H+ ions of 1.00728 amu, 13.6 eV ionisation)."""
    if gas == pic2d.GAS_HYDROGEN:
        p.ion_mass = H_MASS
        p.threshold = 13.6
        p.vti = math.sqrt(p.vti ** 2 * AR_MASS / H_MASS)   # same ion temperature
        p.sigma_load_i = p.vti
    with_argon(p, pressure_pa, temperature)
    p.gas = gas
    return p


def ensemble(nruns=64, seed0=1000):
    """The C4 ensemble: run r seeded seed0 + r; cells in {32^2, 64^2, 128^2,
    256^2}; ppc ~ U[20, 200]; pressure ~ logU[10, 500] mTorr; Te ~ U[1, 5] eV;
    e + Ar+ or e + H+ with probability 1/2 each."""
    import numpy as np
    runs = []
    for r in range(nruns):
        g = np.random.default_rng(seed0 + r)
        cells = int(g.choice([32, 64, 128, 256]))
        ppc = int(g.integers(20, 201))
        mtorr = float(np.exp(g.uniform(np.log(10.0), np.log(500.0))))
        te = float(g.uniform(1.0, 5.0))
        gas = pic2d.GAS_ARGON if g.random() < 0.5 else pic2d.GAS_HYDROGEN
        p = discharge(cells, ppc, te=te, seed=seed0 + r)
        with_gas(p, gas, mtorr * 0.13332)
        runs.append(p)
    return runs


def serving_requests(rate, window_s, seed=7, step_scale=0.1, max_requests=None):
    """C5 request stream: Poisson arrivals of `rate` requests/s over window_s;
    cells in {64^2, 128^2, 256^2, 512^2}; 0.1-20 M particles per species
    (log-uniform, 10-400 ppc); argon or hydrogen as in `ensemble`;
    steps ~ U[500, 5000] * step_scale; multigrid to a relative residual of
    1e-9.  Returns [(Params, steps, arrival_s)]."""
    import numpy as np
    g = np.random.default_rng(seed)
    out = []
    t = float(g.exponential(1.0 / rate))
    k = 0
    while t < window_s and (max_requests is None or k < max_requests):
        cells = int(g.choice([64, 128, 256, 512]))
        n = float(np.exp(g.uniform(np.log(1e5), np.log(2e7))))
        ppc = int(min(400, max(10, round(n / (cells * cells)))))
        mtorr = float(np.exp(g.uniform(np.log(10.0), np.log(500.0))))
        te = float(g.uniform(1.0, 5.0))
        gas = pic2d.GAS_ARGON if g.random() < 0.5 else pic2d.GAS_HYDROGEN
        steps = max(1, int(round(g.uniform(500, 5000) * step_scale)))
        p = discharge(cells, ppc, te=te, seed=2000 + k)
        with_gas(p, gas, mtorr * 0.13332)
        # the request's phi is held within 1e-9 of the FP64 solve: at a
        # relative residual of 1e-8 the error of some requests reached 1e-9
        p.mg_tol = 1e-9
        out.append((p, steps, t))
        t += float(g.exponential(1.0 / rate))
        k += 1
    return out


def argon_256(**kw):
    """256x256 cells, 200 ppc (13.1 M particles per species), argon 100 mTorr."""
    return with_argon(discharge(256, 200, **kw))


def argon_1024(**kw):
    """1024x1024 cells, 100 ppc (104.9 M particles per species), argon 100 mTorr."""
    return with_argon(discharge(1024, 100, **kw))


def large_1024(**kw):
    """1024x1024 cells, 100 ppc (104.9 M particles per species)."""
    return discharge(1024, 100, **kw)
