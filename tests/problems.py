"""Problem set-ups restated from Raven.jl's integration tests (test infrastructure).

Each function cites the reference test it follows.  Used by the oracle tests
and the GPU parity tests alike, so both paths see identical inputs.
"""
from fractions import Fraction

import numpy as np


def julia_range(a, b, n):
    """range(FT(a), stop = FT(b), length = n): correctly rounded lattice."""
    a, b = Fraction(a), Fraction(b)
    return np.array([float(a + (b - a) * i / (n - 1)) for i in range(n)])


def julia_step(a, b, n):
    return float((Fraction(b) - Fraction(a)) / (n - 1))


def coord_sum(P, dim):
    s = P[:, 0, :].copy()
    for d in range(1, dim):
        s = s + P[:, d, :]
    return s


def advection_wave(P, dim, t, u=None):
    """test/balancelaws/advection/wave_{1,2,3}d.jl:18-21, u = (1, ..., 1)."""
    usum = float(dim) if u is None else float(sum(u))
    return (2 + np.sin(np.pi * (coord_sum(P, dim) - usum * t))).reshape(-1).copy()


def euler_energy(rho, mom, p, gamma):
    """energy(law, ρ, ρu⃗, p) = p/(γ-1) + ρu⃗'ρu⃗/2ρ (euler.jl:38-41)."""
    dot = mom[0] * mom[0]
    for m in mom[1:]:
        dot = dot + m * m
    return p / (gamma - 1) + dot / (2 * rho)


def euler_wave(P, dim, t, gamma=1.4):
    """test/balancelaws/euler/wave_{1,2,3}d.jl:18-26: ρ-wave, u⃗ = 1, p = 1."""
    rho = 2 + np.sin(np.pi * (coord_sum(P, dim) - dim * t))
    mom = [rho * 1.0 for _ in range(dim)]
    rhoe = euler_energy(rho, mom, 1.0, gamma)
    return np.ascontiguousarray(np.stack([rho] + mom + [rhoe], axis=1)).reshape(-1)


def isentropic_vortex_2d(P, t, gamma=1.4):
    """test/balancelaws/euler/isentropicvortex.jl:18-34."""
    x = P[:, 0, :] - 5 - 1.0 * t
    y = P[:, 1, :] - 0 - 0.0 * t
    f = 5 * np.exp(1 - (x * x + y * y))
    rho = (1 - (gamma - 1) * f**2 / (16 * gamma * np.pi**2)) ** (1 / (gamma - 1))
    u = 1 + f / (2 * np.pi) * (-y)
    v = 0 + f / (2 * np.pi) * x
    mom = [rho * u, rho * v]
    rhoe = euler_energy(rho, mom, rho**gamma, gamma)
    return np.ascontiguousarray(np.stack([rho] + mom + [rhoe], axis=1)).reshape(-1)


def isentropic_vortex_3d(P, t, gamma=1.4, centre=(5.0, 5.0), noise=None):
    """The 2-D vortex above extruded in z (w = 0), centred at `centre`, with an
    optional additive density perturbation `noise` (same shape as P[:, 0, :])."""
    x = P[:, 0, :] - centre[0] - 1.0 * t
    y = P[:, 1, :] - centre[1]
    f = 5 * np.exp(1 - (x * x + y * y))
    rho = (1 - (gamma - 1) * f**2 / (16 * gamma * np.pi**2)) ** (1 / (gamma - 1))
    u = 1 + f / (2 * np.pi) * (-y)
    v = f / (2 * np.pi) * x
    w = np.zeros_like(u)
    p = rho**gamma
    if noise is not None:
        rho = rho + noise
    mom = [rho * u, rho * v, rho * w]
    rhoe = euler_energy(rho, mom, p, gamma)
    return np.ascontiguousarray(np.stack([rho] + mom + [rhoe], axis=1)).reshape(-1)


# Golden cases of Raven's tests: (name, dim, law, domain, cfl, dt rule, T)
def entropy_square_setup(law, N=4, K=5, gamma=1.4):
    """test/balancelaws/{euler,advection}/entropy_conservation_1d.jl: a square
    density pulse (ρ = 2 for |x| < 1/2, else 1) on the periodic brick [-1, 1]
    with K cells of degree N, run to T = 5 by RLSRK54.  Euler: u = 1, p = 1,
    EntropyConservativeFlux in volume and on faces, cfl 1/2 over the sound
    speed; advection: u = 1, CentralFlux in both, cfl 1/4."""
    nq = N + 1
    v = julia_range(-1, 1, K + 1)

    def rho_of(P):
        return np.where(np.abs(P[:, 0, :]) < 0.5, 2.0, 1.0)

    if law == "euler":
        dt = 0.5 * julia_step(-1, 1, K + 1) / nq / np.sqrt(gamma * 1.0 / 1.0)

        def ic(P):
            rho = rho_of(P)
            mom = [rho * 1.0]
            rhoe = euler_energy(rho, mom, 1.0, gamma)
            return np.ascontiguousarray(np.stack([rho] + mom + [rhoe], axis=1)).reshape(-1)
        return dict(nq=nq, coords=[v], law="euler", u=None, flux="ec", ic=ic, dt=dt, T=5.0)
    dt = 0.25 * julia_step(-1, 1, K + 1) / nq / 1.0
    return dict(nq=nq, coords=[v], law="advection", u=[1.0], flux="central",
                ic=lambda P: rho_of(P).reshape(-1).copy(), dt=dt, T=5.0)


def golden_setup(case, K):
    """Grid extents, law, initial condition, dt and end time of a golden case."""
    N = 4
    if case.startswith("advection_"):
        dim = int(case[-2])
        v = julia_range(-1, 1, K + 1)
        cfl = 0.5 if dim == 1 else 0.25
        dt = cfl * julia_step(-1, 1, K + 1) / N / np.sqrt(float(dim))
        return dict(dim=dim, coords=[v] * dim, law="advection", u=[1.0] * dim,
                    ic=lambda P, t: advection_wave(P, dim, t), dt=dt, T=0.7)
    if case.startswith("euler_wave_"):
        dim = int(case[len("euler_wave_")])
        v = julia_range(-1, 1, K + 1)
        cfl = 0.5 if dim == 1 else 0.25
        dt = cfl * julia_step(-1, 1, K + 1) / N / np.sqrt(1.4 * 1.0 / 1.0)
        T = dt if case.endswith("onestep") else 0.7
        return dict(dim=dim, coords=[v] * dim, law="euler", u=None,
                    ic=lambda P, t: euler_wave(P, dim, t), dt=dt, T=T)
    if case == "euler_vortex_2d":
        vx, vy = julia_range(0, 20, K + 1), julia_range(-5, 5, K + 1)
        dt = 0.25 * julia_step(-5, 5, K + 1) / N
        return dict(dim=2, coords=[vx, vy], law="euler", u=None,
                    ic=lambda P, t: isentropic_vortex_2d(P, t), dt=dt, T=5.0)
    raise KeyError(case)


def c3_warp(seed=99):
    """The seeded smooth warp of the adaptive 3-D Euler workload: a periodic
    perturbation of [0,10]^3 with amplitudes in [0.15, 0.3] and random phases.
    Returns (warp, rng), rng advanced past the warp's draws (the workload then
    draws its density noise from it)."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.15, 0.3, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    k = 2 * np.pi / 10

    def warp(x, y, z):
        return (x + amp[0] * np.sin(k * y + ph[0]) * np.sin(k * z),
                y + amp[1] * np.sin(k * x + ph[1]) * np.sin(k * z),
                z + amp[2] * np.sin(k * x) * np.sin(k * y + ph[2]))
    return warp, rng


def c3_gridmanager(R, nq=6, n=16, radius=2.0, levels=2):
    """GridManager of the adaptive workload: an n^3 periodic brick on [0,10]^3
    refined, one level per adapt! call, inside the sphere of `radius` at the
    centre up to `levels` (2:1 balance is part of each call)."""
    gm = R.GridManager(R.LobattoCell(nq, nq, nq),
                       R.brick([np.linspace(0, 10, n + 1)] * 3, (True,) * 3))
    for r in range(levels):
        c = gm.cell_centres()
        flags = np.where((np.linalg.norm(c - 5.0, axis=1) < radius) & (gm.cells()[1] <= r),
                         R.AdaptRefine, R.AdaptNone)
        gm.adapt(flags)
    return gm
