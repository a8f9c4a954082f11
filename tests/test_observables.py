"""Observable estimators on ensembles whose values are known analytically
(CPU only): Wigner samples of the vacuum, a coherent state (g2 = 1) and a
thermal state (g2 = 2)."""
import numpy as np

from pincosimulator_amd import observables
from pincosimulator_amd.engine import Result
from pincosimulator_amd.observables import from_trajectories


def _result(frames, shape):
    s = from_trajectories(frames, shape)
    B, nf, _ = frames.shape
    return Result(n_frames=nf, n_steps=0, t_final=0.0, truncated=False, kernel="test", n_traj=B, shape=shape,
                  **s)


def _wigner(rng, B, n, alpha, nbar):
    # Wigner function of a displaced thermal state: complex Gaussian around
    # alpha with <|delta|^2> = nbar + 1/2
    s = np.sqrt((nbar + 0.5) / 2)
    z = alpha + s * (rng.standard_normal((B, 1, n)) + 1j * rng.standard_normal((B, 1, n)))
    return z


def test_vacuum():
    rng = np.random.default_rng(0)
    res = _result(_wigner(rng, 20000, 16, 0.0, 0.0), (4, 4))
    n, se = observables.mean_occupation(res)
    assert abs(n[0]) < 4 * se[0]
    nk = observables.momentum_occupation(res)
    assert np.max(np.abs(nk)) < 0.05


def test_coherent_state_g2_is_one():
    rng = np.random.default_rng(1)
    res = _result(_wigner(rng, 40000, 8, 1.3 - 0.4j, 0.0), (1, 8))
    g2, se = observables.g2_jackknife(res)
    n, nse = observables.mean_occupation(res)
    assert abs(n[0] - abs(1.3 - 0.4j) ** 2) < 4 * nse[0]
    assert abs(g2[0] - 1.0) < 4 * se[0] and se[0] < 0.02
    assert abs(observables.g2(res)[0] - g2[0]) < 1e-12
    # mean field and k = 0 occupation of a uniform coherent state
    mf = observables.mean_field(res)
    assert np.max(np.abs(mf - (1.3 - 0.4j))) < 0.02
    nk = observables.momentum_occupation(res)[0]
    assert abs(nk[0, 0] - 8 * abs(1.3 - 0.4j) ** 2) < 0.2


def test_thermal_state_g2_is_two():
    rng = np.random.default_rng(2)
    res = _result(_wigner(rng, 60000, 8, 0.0, 1.5), (2, 4))
    g2, se = observables.g2_jackknife(res)
    n, nse = observables.mean_occupation(res)
    assert abs(n[0] - 1.5) < 4 * nse[0]
    assert abs(g2[0] - 2.0) < 4 * se[0] and se[0] < 0.05


def test_jackknife_matches_bootstrap_scale():
    rng = np.random.default_rng(3)
    res = _result(_wigner(rng, 4000, 4, 0.5, 0.5), (1, 4))
    g2, se = observables.g2_jackknife(res)
    boots = []
    m2, m4 = res.traj_m2[:, 0], res.traj_m4[:, 0]
    for _ in range(300):
        i = rng.integers(0, m2.size, m2.size)
        a, b = m2[i].mean(), m4[i].mean()
        boots.append((b - 2 * a + 0.5) / (a - 0.5) ** 2)
    assert 0.7 < se[0] / np.std(boots) < 1.4


def test_site_resolved_parameters_must_match_the_lattice():
    """Site-resolved values and noise amplitudes are checked against the
    lattice size before a pointer to them reaches the C ABI."""
    import pytest
    from pincosimulator_amd.engine import _as_matrix
    keep = []
    m = _as_matrix({"base": np.ones(6), "noise": ("gaussian", np.full(6, 0.1))}, 6, keep)
    assert m.noise_type != 0 and len(keep) == 2
    with pytest.raises(ValueError, match="lattice has 6"):
        _as_matrix(np.ones(5), 6, [])
    with pytest.raises(ValueError, match="noise amplitude has 5 entries"):
        _as_matrix({"base": 1.0, "noise": ("uniform", np.ones(5))}, 6, [])
