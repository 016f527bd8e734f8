"""Generate the Lotka-Volterra observation fixture (lv_obs.json).

du1 = a u1 - b u1 u2, du2 = -c u2 + d u1 u2 with (a, b, c, d) = (1.5, 1, 3, 1),
u0 = (1, 1), t in [0, 10] sampled at 101 points: classical RK4 in FP64 with
dt = 0.01 (10 steps between samples), plus Gaussian noise N(0, 0.05^2) drawn
from numpy's default_rng(20240601).  Run: python tests/golden/make_lv_obs.py

Also lv_obs_benchmark.json, the data of the reference's Lotka-Volterra benchmark
(benchmarks/Lotka_Volterra/lotka_volterra.jl:21-37: the same system on
t in [0, 3] sampled every 0.01, 301 points, noise-free from a tight-tolerance
solve): here FP64 RK4 with 20 steps of 0.0005 between samples (global error
~1e-12).

And fhn_obs_benchmark.json, the data of the reference's FitzHugh-Nagumo benchmark
(benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-38): dv = v - v^3/3 - w + l,
dw = tau_inv (v + a - b w), (a, b, tau_inv, l) = (0.7, 0.8, 0.08, 0.5), u0 = (1, 1),
t in [0, 3] sampled every 0.01 (301 points), noise-free; FP64 RK4 as above.
"""
import json
import os

import numpy as np

THETA = (1.5, 1.0, 3.0, 1.0)
U0 = (1.0, 1.0)
DT = 0.01
STEPS_PER_OBS = 10
N_OBS = 101
SIGMA = 0.05
SEED = 20240601


FHN_THETA = (0.7, 0.8, 0.08, 0.5)


def fhn_rhs(theta):
    a, b, tinv, l = theta
    return lambda u: np.array([u[0] - u[0] ** 3 / 3 - u[1] + l, tinv * (u[0] + a - b * u[1])])


def rk4_trajectory(theta=THETA, u0=U0, dt=DT, spo=STEPS_PER_OBS, n_obs=N_OBS, rhs=None):
    a, b, c, d = theta
    f = rhs or (lambda u: np.array([a * u[0] - b * u[0] * u[1], -c * u[1] + d * u[0] * u[1]]))
    u = np.array(u0, np.float64)
    out = [u.copy()]
    for _ in range(n_obs - 1):
        for _ in range(spo):
            k1 = f(u)
            k2 = f(u + dt / 2 * k1)
            k3 = f(u + dt / 2 * k2)
            k4 = f(u + dt * k3)
            u = u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.array(out)


def main():
    traj = rk4_trajectory()
    noise = np.random.default_rng(SEED).normal(0.0, SIGMA, traj.shape)
    obs = traj + noise
    doc = {"theta": THETA, "u0": U0, "dt": DT, "steps_per_obs": STEPS_PER_OBS, "sigma": SIGMA,
           "seed": SEED, "trajectory": traj.tolist(), "obs": obs.tolist()}
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "lv_obs.json")
    with open(path, "w") as fh:
        json.dump(doc, fh)
    print(f"wrote {path}")
    bench = rk4_trajectory(dt=0.0005, spo=20, n_obs=301)
    doc = {"theta": THETA, "u0": U0, "saveat": 0.01, "t_end": 3.0, "integrator":
           "FP64 RK4, dt = 0.0005", "obs": bench.tolist()}
    path = os.path.join(here, "lv_obs_benchmark.json")
    with open(path, "w") as fh:
        json.dump(doc, fh)
    print(f"wrote {path}")
    fhn = rk4_trajectory(FHN_THETA, (1.0, 1.0), 0.0005, 20, 301, rhs=fhn_rhs(FHN_THETA))
    doc = {"theta": FHN_THETA, "u0": (1.0, 1.0), "saveat": 0.01, "t_end": 3.0,
           "integrator": "FP64 RK4, dt = 0.0005", "obs": fhn.tolist()}
    path = os.path.join(here, "fhn_obs_benchmark.json")
    with open(path, "w") as fh:
        json.dump(doc, fh)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
