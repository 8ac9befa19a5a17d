"""Host-side pieces of the package that need no GPU: the binning analysis of
the block energies and the cache of prepared systems."""

import numpy as np
import pytest

from oracle import afqmc as O


@pytest.mark.parametrize("kind", ["white", "ar1", "short", "constant"])
def test_do_binning_analysis_matches_oracle(kind):
    """omegaqmc_amd.do_binning_analysis (utils.py:86-149) against the oracle:
    mean, error, standard deviation and the autocorrelation time kappa."""
    from omegaqmc_amd.afqmc_gto import do_binning_analysis
    rng = np.random.default_rng(4)
    if kind == "white":
        a = rng.normal(size=200)
    elif kind == "ar1":
        a = np.zeros(300)
        for i in range(1, 300):
            a[i] = 0.8 * a[i - 1] + rng.normal()
    elif kind == "short":
        a = rng.normal(size=5) - 100.0
    else:
        a = np.full(20, -1.25)
    got = do_binning_analysis(a)
    if kind == "constant":
        # 0/0 autocorrelation: no positive lag, kappa = 1, error 0
        assert got[0] == -1.25 and got[1] == 0.0 and got[2] == 0.0 and got[3] == 1.0
        return
    ref = O.do_binning_analysis(a)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)
    if kind == "ar1":
        # kappa of an AR(1) series with phi = 0.8 is (1 + phi) / (1 - phi) = 9
        # (estimate truncated at the first negative lag)
        assert 4.0 < got[3] < 14.0


def test_system_cache_keeps_inputs_and_is_bounded():
    """Entries hold the caller's arrays (their ids cannot be reused while the
    entry lives); a fresh array never hits an old entry; the cache keeps at
    most SYSTEM_CACHE_SIZE entries."""
    from omegaqmc_amd import afqmc_gto as A
    prop = {}
    made = []

    def make():
        made.append(object())
        return made[-1]

    a, b = np.zeros(3), np.zeros(3)
    s1 = A._cached_system(prop, (a, b), ("fp64",), make)
    assert A._cached_system(prop, (a, b), ("fp64",), make) is s1
    assert A._cached_system(prop, (a, b), ("bf16x3",), make) is not s1
    for _ in range(3 * A.SYSTEM_CACHE_SIZE):
        # temporaries: freed right after the call unless the cache holds them
        A._cached_system(prop, (np.zeros(3), b), ("fp64",), make)
    assert len(prop['_systems']) == A.SYSTEM_CACHE_SIZE
    held = [x for k, (arrs, _) in prop['_systems'].items() for x in arrs]
    assert all(isinstance(x, np.ndarray) for x in held)
    assert len(made) == 2 + 3 * A.SYSTEM_CACHE_SIZE
