"""Synthetic observation tracks for the constant-velocity PF workloads.

The reference reads its observations from example/pf/pf.data; the workloads
here simulate them once from a seed (numpy PCG64): a 4-D constant-velocity
state (x, y, vx, vy) with dt = 0.1 and Gaussian process noise, observed in
position with Student-t noise (df 10, scale 0.1) or Gaussian noise, with an
optional fraction of outliers at a multiple of the scale.
"""
import numpy as np


def cv_simulate(T, seed, dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=2.0, sd_vel0=1.0,
                obs="student", obs_scale=0.1, obs_df=10.0, outlier_frac=0.0,
                outlier_mult=10.0):
    """Returns (states[T, 4], obs_x[T], obs_y[T])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.array([rng.normal(0, sd_pos0), rng.normal(0, sd_pos0),
                  rng.normal(0, sd_vel0), rng.normal(0, sd_vel0)])
    states = np.zeros((T, 4))
    ox = np.zeros(T)
    oy = np.zeros(T)
    for t in range(T):
        if t > 0:
            x = np.array([x[0] + rng.normal(0, sd_pos) + dt * x[2],
                          x[1] + rng.normal(0, sd_pos) + dt * x[3],
                          x[2] + rng.normal(0, sd_vel),
                          x[3] + rng.normal(0, sd_vel)])
        states[t] = x
        if obs == "student":
            e = rng.standard_t(obs_df, size=2) * obs_scale
        else:
            e = rng.normal(0, obs_scale, size=2)
        if outlier_frac > 0 and rng.random() < outlier_frac:
            e = e * outlier_mult
        ox[t] = x[0] + e[0]
        oy[t] = x[1] + e[1]
    return states, ox, oy


def cv_simulate_many(F, T, seed, dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=2.0, sd_vel0=1.0,
                     obs="student", obs_scale=0.1, obs_df=10.0, outlier_frac=0.0,
                     outlier_mult=10.0):
    """F independent tracks of T steps at once (vectorised over tracks):
    returns (obs_x[F, T], obs_y[F, T])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.stack([rng.normal(0, sd_pos0, F), rng.normal(0, sd_pos0, F),
                  rng.normal(0, sd_vel0, F), rng.normal(0, sd_vel0, F)])
    ox = np.zeros((F, T))
    oy = np.zeros((F, T))
    for t in range(T):
        if t > 0:
            x = np.stack([x[0] + rng.normal(0, sd_pos, F) + dt * x[2],
                          x[1] + rng.normal(0, sd_pos, F) + dt * x[3],
                          x[2] + rng.normal(0, sd_vel, F),
                          x[3] + rng.normal(0, sd_vel, F)])
        if obs == "student":
            e = rng.standard_t(obs_df, size=(2, F)) * obs_scale
        else:
            e = rng.normal(0, obs_scale, size=(2, F))
        if outlier_frac > 0:
            e = e * np.where(rng.random(F) < outlier_frac, outlier_mult, 1.0)
        ox[:, t] = x[0] + e[0]
        oy[:, t] = x[1] + e[1]
    return ox, oy


def gmm_simulate(n=1000, seed=2016, means=(-3.0, 0.0, 2.0, 5.0), sds=(0.5, 1.5, 0.75, 1.0),
                 weights=(0.1, 0.2, 0.3, 0.4)):
    """n points from a seeded 1-D Gaussian mixture (the tempered SMC
    workload: 4 components, means -3, 0, 2, 5, sd 0.5-1.5, weights
    0.1-0.4)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    comp = rng.choice(len(means), size=n, p=np.asarray(weights))
    return rng.normal(np.asarray(means)[comp], np.asarray(sds)[comp])
