"""MMA test problems: the reference optimiser (src/utility/mma.h via oracle/_ref) writes their
fixtures (tests/golden/make_golden.py mma); the GPU optimiser replays them (tests/test_gpu_mma.py).

Each problem runs `steps` consecutive updates of one optimiser.  Step k feeds the design the
reference produced at step k-1 (x0 seeded) with seeded objective gradients and linear constraints
evaluated at that design, so every step is compared from the same inputs and the optimiser's
history (asymptotes, the previous two designs) is the reference's.  The problems cover the heat
sink's setting (one volume constraint, a0 = 1, a = 0, c = 1e4, d = 0, heatsink3D.cpp:78-85) and
two constraints with d > 0; the steps cover the asymptote initialisation (k < 2) and the
oscillation / monotone / neutral branches after it (mma.h:198-224).
"""
import numpy as np

PROBLEMS = [
    dict(name="mma_n400_m1", n=400, m=1, a0=1.0, a=[0.0], c=[1e4], d=[0.0], move=0.2, steps=6, seed=5),
    dict(name="mma_n300_m2", n=300, m=2, a0=1.0, a=[0.0, 0.0], c=[1e3, 1e3], d=[1.0, 1.0], move=0.5, steps=5,
         seed=6),
]


def x0(pr):
    return np.random.default_rng(pr["seed"]).uniform(0.3, 0.7, pr["n"])


def step_inputs(pr, k, x):
    """(f, dfdx, g, dgdx) of step k at design x."""
    n, m = pr["n"], pr["m"]
    rng = np.random.default_rng(1000 * pr["seed"] + k)
    w = rng.uniform(0.5, 1.5, n)
    dfdx = -w * (1.2 - x) + 0.3 * rng.standard_normal(n)  # mixed signs: both p0/q0 branches
    f = float(np.sum(w * (1.2 - x) ** 2) / 2.0)
    rows = [np.full(n, 1.0 / n)]
    g = [float(x.mean() - 0.4)]
    if m == 2:
        r = rng.uniform(-1.0, 1.0, n) / n
        rows.append(r)
        g.append(float(r @ x - 0.01))
    return f, dfdx, np.array(g), np.array(rows)


def xbounds(pr):
    return np.zeros(pr["n"]), np.ones(pr["n"])
