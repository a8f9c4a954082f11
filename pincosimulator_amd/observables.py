"""Ensemble observables of a TWMC run, from the engine's FP64 sums.

The reference leaves observables to an external Matlab package, so they are
defined here, in Wigner (symmetric) ordering with the unnormalised forward DFT
beta~_k = sum_r beta_r exp(-i k.r):

    n_i     = <|beta_i|^2> - 1/2
    n_k     = <|beta~_k|^2> / N - 1/2
    g2(0)   = mean_i(<|beta_i|^4> - 2 <|beta_i|^2> + 1/2) / (mean_i <|beta_i|^2> - 1/2)^2

Standard errors are over trajectories: <n> from the per-trajectory site means,
g2(0) by the leave-one-out jackknife on the per-trajectory site means of
|beta|^2 and |beta|^4 (site-averaged g2 is a function of those two means only).
"""
from __future__ import annotations

import numpy as np


def site_occupation(res):
    """n_i per frame, shape (frames, nx, ny)."""
    return (res.sum_n / res.n_traj - 0.5).reshape((-1,) + tuple(res.shape))


def momentum_occupation(res):
    """n_k per frame, shape (frames, nx, ny) (k in FFT order)."""
    n = res.shape[0] * res.shape[1]
    return (res.sum_nk / res.n_traj / n - 0.5).reshape((-1,) + tuple(res.shape))


def mean_field(res):
    return (res.sum_beta / res.n_traj).reshape((-1,) + tuple(res.shape))


def mean_occupation(res):
    """(<n>, standard error) per frame, averaged over sites."""
    m2 = res.traj_m2  # (B, frames)
    B = m2.shape[0]
    mean = m2.mean(axis=0) - 0.5
    se = m2.std(axis=0, ddof=1) / np.sqrt(B) if B > 1 else np.zeros_like(mean)
    return mean, se


def _g2_from_means(m2, m4):
    return (m4 - 2.0 * m2 + 0.5) / (m2 - 0.5) ** 2


def g2(res):
    """Site-averaged g2(0) per frame from the ensemble sums."""
    B = res.n_traj
    m2 = (res.sum_n / B).mean(axis=1)
    m4 = (res.sum_n2 / B).mean(axis=1)
    return _g2_from_means(m2, m4)


def g2_jackknife(res):
    """(g2, jackknife standard error) per frame from per-trajectory site means."""
    m2b, m4b = res.traj_m2, res.traj_m4
    B = m2b.shape[0]
    s2, s4 = m2b.sum(axis=0), m4b.sum(axis=0)
    full = _g2_from_means(s2 / B, s4 / B)
    if B < 2:
        return full, np.zeros_like(full)
    loo = _g2_from_means((s2[None, :] - m2b) / (B - 1), (s4[None, :] - m4b) / (B - 1))
    se = np.sqrt((B - 1) / B * ((loo - loo.mean(axis=0)) ** 2).sum(axis=0))
    return full, se


def from_trajectories(frames, shape):
    """Same estimators from raw trajectory frames (B, frames, N) complex --
    used to compare engine sums with explicitly stored trajectories."""
    frames = np.asarray(frames)
    B = frames.shape[0]
    n = frames.shape[2]
    a2 = np.abs(frames) ** 2
    fk = np.fft.fft2(frames.reshape(B, frames.shape[1], *shape), axes=(-2, -1)).reshape(B, frames.shape[1], n)
    return {
        "sum_n": a2.sum(axis=0), "sum_n2": (a2 ** 2).sum(axis=0),
        "sum_nk": (np.abs(fk) ** 2).sum(axis=0), "sum_beta": frames.sum(axis=0),
        "traj_m2": a2.mean(axis=2), "traj_m4": (a2 ** 2).mean(axis=2),
    }
