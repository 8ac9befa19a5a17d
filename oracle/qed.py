"""Cavity-QED phaseless AFQMC, NumPy FP64 restatement (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows OmegaQMC/qed_afqmc_gto.py (propagate_photon 262-311,
_apply_exp_vhs_photon 314-332, propagate_qed_walkers 335-440,
qed_local_energy 156-201, population_control_comb_qed 106-153,
build_qed_propagator 204-259) and the array part of
OmegaQMC/integrals/qed.py:95-205 (prepare_qed_integrals).  As in the
other oracles the random draws are arguments: ``xi`` (W, G+1) for the
auxiliary fields, ``eta1``/``eta2`` (W,) for the two photon half steps.

Pins (tests/test_oracle_qed.py):
  propagate_photon   the sampled density times the weight factor equals the
                     importance-sampled harmonic-oscillator propagator
                     f(q') K(q, q') / f(q) with K Mehler's closed form (zero
                     point subtracted), for any trial (s, q0).
  qed_local_energy   the photon term is <f|H_ph|q>/<f|q> by finite
                     differences of f; with zero coupling it reduces to the
                     electronic local energy plus the photon term.
  apply_exp_vhs_photon  scipy expm within the Taylor remainder.
  propagate_qed_walkers  with zero coupling the electrons follow the plain
                     step (the dipole vector is zero) and only the photon
                     weight factor multiplies the weights.
"""

from __future__ import annotations

import numpy as np

from . import afqmc as O


def prepare_qed_integrals_from_arrays(h1e, chol_mo, dip_mo, quad_mo=None):
    """integrals/qed.py:95-205 for MO-basis arrays: h1e (+ 0.5 quad_mo, the
    proper dipole self-energy), chol_qed = [chol; dip], v0 and h1e_mod_0."""
    h1e = np.asarray(h1e, dtype=np.float64)
    if quad_mo is not None:
        h1e = h1e + 0.5 * np.asarray(quad_mo)
    chol_qed = np.concatenate([np.asarray(chol_mo), np.asarray(dip_mo)[None]], axis=0)
    G, M = chol_qed.shape[0], chol_qed.shape[1]
    A = chol_qed.transpose(0, 2, 1).reshape(G * M, M)
    v0 = -0.5 * (A.T @ A)
    return {'h1e': h1e, 'h1e_mod_0': h1e + v0, 'chol_qed': chol_qed, 'dip_mo': np.asarray(dip_mo)}


def build_qed_propagator(h1e_mod_0, chol_qed, trial_up, trial_dn, dt, G_charge=None):
    """qed_afqmc_gto.py:204-259 -> {'expH1_0', 'mf_shift', 'dt'} (the plain
    propagator's formulas on the augmented vectors)."""
    p = O.build_propagator(h1e_mod_0, chol_qed, trial_up, trial_dn, dt, G_charge)
    return {'expH1_0': p['expH1'], 'mf_shift': p['mf_shift'], 'dt': dt}


def propagate_photon(q, omega, dt, s, q0, eta):
    """qed_afqmc_gto.py:262-311 with the normal draws eta given -> (q_new, log_w)."""
    delta = omega * dt / 2.0
    cosh_d = np.cosh(delta)
    tanh_d = np.tanh(delta)
    drift = -s * (q - q0)
    q_mean = q / cosh_d + tanh_d * drift
    q_new = q_mean + np.sqrt(tanh_d) * eta
    log_trial_ratio = -s / 2.0 * ((q_new - q0) ** 2 - (q - q0) ** 2)
    correction = (q / cosh_d - q_new) * drift
    additional = -tanh_d / 2.0 * (-(drift ** 2) + q ** 2) + delta / 2.0
    return q_new, log_trial_ratio + correction + additional - 0.5 * np.log(cosh_d)


def apply_exp_vhs_photon(coeff, mat, phi, nmax=4):
    """qed_afqmc_gto.py:314-332 (no chunking): exp(coeff[w] mat) phi[w] by Taylor."""
    VHS = np.asarray(coeff)[:, None, None] * np.asarray(mat)[None]
    return O.apply_exp_vhs(VHS, phi, nmax)


def qed_local_energy(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc, q, dip_mo, omega, s,
                     q0):
    """qed_afqmc_gto.py:156-201 -> (e_tot, e_1b, e_2b, e_ph)."""
    e_1b_0 = O.local_energy_1body(h1e, Ga, Gb, enuc)
    tr = np.einsum('pq,wqp->w', dip_mo, Ga) + np.einsum('pq,wqp->w', dip_mo, Gb)
    e_ep = np.sqrt(omega) * q * tr
    e_coul, e_exch = O.local_energy_2body(Ghalfa, Ghalfb, rchol_a, rchol_b)
    e_2b = e_coul - e_exch
    e_ph = omega / 2.0 * (s - s ** 2 * (q - q0) ** 2 + q ** 2 - 1.0)
    e_1b = e_1b_0 + e_ep
    return e_1b + e_2b + e_ph, e_1b, e_2b, e_ph


def propagate_qed_walkers(phia, phib, q, weights, overlap, e_hybrid, propagator, chol_qed,
                          rchol_a, rchol_b, trial_up, trial_dn, eshift, xi, eta1, eta2, dip_mo,
                          omega, s, q0, fbbound=None, exp_nmax=6):
    """qed_afqmc_gto.py:335-440 with the draws given
    -> (phia, phib, q, weights_new, overlap_new, e_hybrid_new)."""
    dt = propagator['dt']
    expH1_0 = propagator['expH1_0']
    mf_shift = propagator['mf_shift']
    Gha, Ghb, ovlp = O.greens_function_force_bias(phia, phib, trial_up, trial_dn)
    q, lw1 = propagate_photon(q, omega, dt, s, q0, eta1)
    c1 = -0.5 * dt * np.sqrt(omega) * q
    phia = apply_exp_vhs_photon(c1, dip_mo, phia, 4)
    phib = apply_exp_vhs_photon(c1, dip_mo, phib, 4)
    phia = np.matmul(expH1_0, phia)
    phib = np.matmul(expH1_0, phib)
    _, xshifted, cfb, cmf, _ = O.force_bias_shift(Gha, Ghb, rchol_a, rchol_b, mf_shift, xi, dt,
                                                  fbbound)
    VHS = O.build_vhs(xshifted, chol_qed, dt)
    phia = O.apply_exp_vhs(VHS, phia, exp_nmax)
    phib = O.apply_exp_vhs(VHS, phib, exp_nmax)
    phia = np.matmul(expH1_0, phia)
    phib = np.matmul(expH1_0, phib)
    c2 = -0.5 * dt * np.sqrt(omega) * q
    phia = apply_exp_vhs_photon(c2, dip_mo, phia, 4)
    phib = apply_exp_vhs_photon(c2, dip_mo, phib, 4)
    q, lw2 = propagate_photon(q, omega, dt, s, q0, eta2)
    ovlp_new = O.greens_function_overlap(phia, phib, trial_up, trial_dn)
    w_new, e_new = O.update_weights_phaseless(weights, ovlp, ovlp_new, cfb, cmf, e_hybrid,
                                              eshift, dt)
    w_new = w_new * np.exp(np.clip(lw1 + lw2, -10.0, 10.0))
    return phia, phib, q, w_new, ovlp_new, e_new


def population_control_comb_qed(weights, phia, phib, q, r, e_hybrid=None):
    """qed_afqmc_gto.py:106-153 with the offset r given
    -> (weights_new, phia, phib, q, e_hybrid, idx)."""
    idx = O.comb_indices(weights, r)
    wn = np.ones_like(np.asarray(weights, dtype=np.float64))
    eh = None if e_hybrid is None else e_hybrid[idx]
    return wn, phia[idx], phib[idx], q[idx], eh, idx
