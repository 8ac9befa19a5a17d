"""Electron-gas (plane-wave, symmetrised Cholesky) step and local energy,
NumPy FP64 restatement (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``OmegaQMC/afqmc_pw_heg.py``: the 3DEG step
(``_propagate_3deg_walkers``, afqmc_pw_heg.py:833-945) with the auxiliary
fields ``xi`` given, and the sign-weighted local energy
(``local_energy_2body_pw`` / ``local_energy_pw``, afqmc_pw_heg.py:414-467).
The Green's functions, the Taylor exponential and the weight update are
those of :mod:`oracle.afqmc` (the same reference functions).

Pins (tests/test_heg.py):
  local_energy_2body_pw  the Fermi-sea determinant gives zero Coulomb (every
                      L^q with q != 0 is off-diagonal in k) and the exchange
                      1/2 sum_s sum_{k != k' occupied} 4 pi / (V |k - k'|^2)
                      summed directly over plane-wave pairs; with all signs
                      +1 it equals oracle.afqmc.local_energy_2body.
  propagate_3deg_walkers  all signs +1 (phase i) reproduces
                      oracle.afqmc.propagate_walkers to rounding (the VHS
                      sum runs in another order); with only antisymmetric
                      (sign -1, phase 1) vectors and real fields the VHS is
                      real antisymmetric, so exp(VHS) keeps phi^H phi (and
                      phase i, which makes it Hermitian, does not); with no vectors
                      the walkers move by expH1 alone and the local energy
                      of the Fermi sea stays the sum of its kinetic energies.
"""

import numpy as np

from . import afqmc as A


def local_energy_2body_pw(Ghalfa, Ghalfb, rchol_a, rchol_b, chol_sign):
    """afqmc_pw_heg.py:414-457 -> (e_coul, e_exch), both weighted by the ERI
    sign of each vector."""
    Xa = np.einsum('giq,wiq->gw', rchol_a, Ghalfa)
    Xb = np.einsum('giq,wiq->gw', rchol_b, Ghalfb)
    e_coul = 0.5 * np.einsum('g,gw->w', chol_sign, (Xa + Xb) ** 2)
    Ta = np.einsum('giq,wjq->gwij', rchol_a, Ghalfa)
    Tb = np.einsum('giq,wjq->gwij', rchol_b, Ghalfb)
    ea = np.sum(Ta * Ta.transpose(0, 1, 3, 2), axis=(2, 3))
    eb = np.sum(Tb * Tb.transpose(0, 1, 3, 2), axis=(2, 3))
    e_exch = 0.5 * np.einsum('g,gw->w', chol_sign, ea + eb)
    return e_coul, e_exch


def local_energy_pw(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc, chol_sign):
    """afqmc_pw_heg.py:460-467 -> (e_tot, e_1b, e_2b)."""
    e_1b = A.local_energy_1body(h1e, Ga, Gb, enuc)
    e_coul, e_exch = local_energy_2body_pw(Ghalfa, Ghalfb, rchol_a, rchol_b, chol_sign)
    e_2b = e_coul - e_exch
    return e_1b + e_2b, e_1b, e_2b


def propagate_3deg_walkers(phia, phib, weights, overlap, e_hybrid, propagator, chol, rchol_a,
                           rchol_b, trial_up, trial_dn, eshift, xi, has_coulomb, ndown=None,
                           fbbound=None, exp_nmax=6, chol_phase=None):
    """afqmc_pw_heg.py:833-945 with the auxiliary fields ``xi`` (W, G) given.

    chol_phase[g] is 1j for sign +1 and 1 for sign -1 (None: 1j for all);
    ndown == 0 leaves the beta walker (the reference's one-column dummy)
    untouched.  Returns (phia, phib, weights_new, overlap_new, e_hybrid_new).
    """
    dt = propagator['dt']
    expH1 = propagator['expH1']
    mf_shift = propagator['mf_shift']
    W = phia.shape[0]
    naux = chol.shape[0]
    skip_beta = ndown is not None and ndown == 0
    if fbbound is None:
        fbbound = A.FBBOUND_DEFAULT
    Gha, Ghb, ovlp = A.greens_function_force_bias(phia, phib, trial_up, trial_dn)
    phia = np.matmul(expH1, phia)
    if not skip_beta:
        phib = np.matmul(expH1, phib)
    if has_coulomb and naux > 0:
        vbias = np.einsum('giq,wiq->gw', rchol_a, Gha)
        if not skip_beta:
            vbias = vbias + np.einsum('giq,wiq->gw', rchol_b, Ghb)
        phase = 1j * np.ones(naux) if chol_phase is None else np.asarray(chol_phase)
        xbar = (-np.sqrt(dt) * (phase[:, None] * vbias - mf_shift[:, None])).T
        xabs = np.abs(xbar)
        with np.errstate(invalid='ignore', divide='ignore'):
            xbar = np.where(xabs > fbbound, xbar * fbbound / xabs, xbar)
        xshifted = xi - xbar
        cmf = -np.sqrt(dt) * np.einsum('wg,g->w', xshifted, mf_shift)
        cfb = np.sum(xi * xbar, axis=1) - 0.5 * np.sum(xbar * xbar, axis=1)
        VHS = np.sqrt(dt) * np.einsum('wg,g,gpq->wpq', xshifted, phase, chol)
        phia = A.apply_exp_vhs(VHS, phia, exp_nmax)
        if not skip_beta:
            phib = A.apply_exp_vhs(VHS, phib, exp_nmax)
    else:
        cmf = np.zeros(W)
        cfb = np.zeros(W)
    phia = np.matmul(expH1, phia)
    if not skip_beta:
        phib = np.matmul(expH1, phib)
    ovlp_new = A.greens_function_overlap(phia, phib, trial_up, trial_dn)
    w_new, e_new = A.update_weights_phaseless(weights, ovlp, ovlp_new, cfb, cmf, e_hybrid,
                                              eshift, dt)
    return phia, phib, w_new, ovlp_new, e_new
