"""Shared inputs of the parity tests (NumPy, host)."""

from __future__ import annotations

import numpy as np


def make_harness(seed=0, nwalkers=8, nbasis=10, nup=3, ndown=3, naux=24):
    """The reference's random harness (tests/test_afqmc_kernels.py:155-177), same
    draw order, so the arrays equal the ones the reference's own tests use."""
    rng = np.random.default_rng(seed)

    def _crand(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    trial_up = _crand((nbasis, nup))
    trial_dn = _crand((nbasis, ndown))
    phia = _crand((nwalkers, nbasis, nup))
    phib = _crand((nwalkers, nbasis, ndown))
    rchol_a = _crand((naux, nup, nbasis))
    rchol_b = _crand((naux, ndown, nbasis))
    h1e = rng.standard_normal((nbasis, nbasis)) + 1j * rng.standard_normal((nbasis, nbasis))
    h1e = 0.5 * (h1e + h1e.conj().T)
    enuc = float(rng.standard_normal())
    return trial_up, trial_dn, phia, phib, rchol_a, rchol_b, h1e, enuc


def walkers_near_trial(trial, nwalkers, rng, scale=0.05):
    """Random complex walkers around a trial (test_afqmc_kernels.py:75-94 recipe)."""
    M, N = trial.shape
    return (np.repeat(trial[None].astype(np.complex128), nwalkers, axis=0)
            + scale * (rng.normal(size=(nwalkers, M, N)) + 1j * rng.normal(size=(nwalkers, M, N))))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
