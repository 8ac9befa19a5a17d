"""Multi-determinant trial (occupation-string dets), NumPy FP64 restatement (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows OmegaQMC/observables/greens_th4.py, observables/energy_th4.py, the
streamed pair of observables/greens.py:314-369 and energy.py:318-395, and
afqmc_gto.py:728-839 (propagate_walkers_multidet); the trial set-up follows
afqmc_gto.py:1163-1225.  As in the single-determinant oracle, the auxiliary
fields ``xi`` are an argument instead of a jax.random draw.

Pins (tests/test_oracle_multidet.py):
  greens_function_multidet*   overlap equals sum_I c_I^* det(T_I^H phi_a)
                      det(T_I^H phi_b) evaluated by Cauchy-Binet on CI vectors
                      (no slogdet, no inverse); one determinant reduces to the
                      single-determinant functions; G of an exact-trial
                      expansion; masking of exactly orthogonal dets.
  local_energy_multidet  equals <Psi_T|H|phi>/<Psi_T|phi> from the
                      brute-force second-quantised Hamiltonian; for Psi_T the
                      exact ground state (all dets, FCI coefficients) it is
                      E_0 for every walker (zero variance).
  propagate_walkers_multidet  one determinant reproduces the single-det step.
  local_energy_multidet_streamed (with greens_function_multidet_force_bias)
                      equals local_energy_multidet (full G, chunked dets)
                      for walkers whose dets are all above the mask.
"""

from __future__ import annotations

import numpy as np

from . import afqmc as O

OVLP_MASK_RTOL = 1e-14   # greens_th4.py:24
OVLP_MASK_ATOL = 1e-300  # greens_th4.py:25


def trials_from_occ(nbasis, occ):
    """afqmc_gto.py:1172-1174: eye[:, occ].transpose(1, 0, 2) -> (ndet, M, N)."""
    eye = np.eye(nbasis)
    occ = np.asarray(occ, dtype=np.int64)
    return eye[:, occ].transpose(1, 0, 2)


def g_charge(trials_up, trials_dn, ci_coeffs):
    """afqmc_gto.py:1209-1218: |c|^2-weighted trial density (spin-summed)."""
    w = np.abs(ci_coeffs) ** 2
    w = w / np.sum(w)
    G = np.einsum('d,dpi,dqi->pq', w, trials_up, trials_up.conj())
    return G + np.einsum('d,dpi,dqi->pq', w, trials_dn, trials_dn.conj())


def _gf_spin_single_det(phi, trial):
    """greens.py:201-222 -> (Ghalf, ovlp); a singular S gives inf/NaN Ghalf
    (jnp.linalg.inv) and ovlp 0 (slogdet sign 0), walker by walker."""
    S = np.einsum('wpi,pj->wij', phi, trial.conj())
    W, N = S.shape[0], S.shape[1]
    Sinv = np.empty_like(S)
    for w in range(W):
        try:
            Sinv[w] = np.linalg.inv(S[w])
        except np.linalg.LinAlgError:
            Sinv[w] = np.nan
    Ghalf = np.einsum('wij,wqj->wiq', Sinv, phi)
    sign, logdet = np.linalg.slogdet(S)
    return Ghalf, sign * np.exp(logdet)


def _overlap_only_spin(phi, trial):
    """greens.py:79-101."""
    S = np.einsum('wpi,pj->wij', phi, trial.conj())
    sign, logdet = np.linalg.slogdet(S)
    return sign * np.exp(logdet)


def det_overlap_max(phia, phib, trials_up, trials_dn):
    """greens_th4.py:27-31."""
    oa = np.stack([_overlap_only_spin(phia, t) for t in trials_up])
    ob = np.stack([_overlap_only_spin(phib, t) for t in trials_dn])
    return np.max(np.abs(oa * ob), axis=0)


def mask_small_overlap(Gha, Ghb, oa, ob, omax):
    """greens_th4.py:34-46 (Gh*: (ndet, W, N, M); o*: (ndet, W))."""
    tol = np.maximum(OVLP_MASK_RTOL * omax, OVLP_MASK_ATOL)
    keep = np.abs(oa * ob) >= tol[None, :]
    k4 = keep[:, :, None, None]
    return np.where(k4, Gha, 0.0), np.where(k4, Ghb, 0.0)


def _chunks(ndet, size):
    return [(d0, min(d0 + size, ndet)) for d0 in range(0, ndet, size)]


def greens_function_multidet(phia, phib, trials_up, trials_dn, ci_coeffs, det_chunk_size=5):
    """greens_th4.py:75-148 -> (Ga, Gb, overlap).

    Zero-padded dets of the reference's last chunk contribute exact zeros and
    are not formed."""
    W, M = phia.shape[0], phia.shape[1]
    omax = det_overlap_max(phia, phib, trials_up, trials_dn)
    Ga = np.zeros((W, M, M), dtype=np.complex128)
    Gb = np.zeros((W, M, M), dtype=np.complex128)
    ovlp = np.zeros(W, dtype=np.complex128)
    for d0, d1 in _chunks(trials_up.shape[0], det_chunk_size):
        Ghs_a, oas, Ghs_b, obs = [], [], [], []
        for d in range(d0, d1):
            gh, o = _gf_spin_single_det(phia, trials_up[d])
            Ghs_a.append(gh)
            oas.append(o)
            gh, o = _gf_spin_single_det(phib, trials_dn[d])
            Ghs_b.append(gh)
            obs.append(o)
        Gha, Ghb = np.stack(Ghs_a), np.stack(Ghs_b)
        oa, ob = np.stack(oas), np.stack(obs)
        Gha, Ghb = mask_small_overlap(Gha, Ghb, oa, ob, omax)
        w = ci_coeffs[d0:d1].conj()[:, None] * oa * ob
        ovlp = ovlp + np.sum(w, axis=0)
        w4 = w[:, :, None, None]
        Ga = Ga + np.sum(w4 * np.einsum('dpi,dwiq->dwpq', trials_up[d0:d1].conj(), Gha), axis=0)
        Gb = Gb + np.sum(w4 * np.einsum('dpi,dwiq->dwpq', trials_dn[d0:d1].conj(), Ghb), axis=0)
    return Ga / ovlp[:, None, None], Gb / ovlp[:, None, None], ovlp


def greens_function_multidet_overlap_only(phia, phib, trials_up, trials_dn, ci_coeffs,
                                          det_chunk_size=5):
    """greens_th4.py:186-228."""
    ovlp = np.zeros(phia.shape[0], dtype=np.complex128)
    for d0, d1 in _chunks(trials_up.shape[0], det_chunk_size):
        oa = np.stack([_overlap_only_spin(phia, trials_up[d]) for d in range(d0, d1)])
        ob = np.stack([_overlap_only_spin(phib, trials_dn[d]) for d in range(d0, d1)])
        w = ci_coeffs[d0:d1].conj()[:, None] * oa * ob
        ovlp = ovlp + np.sum(w, axis=0)
    return ovlp


def local_energy_multidet(h1e, chol, phia, phib, trials_up, trials_dn, ci_coeffs, enuc,
                          det_chunk_size=5):
    """energy_th4.py:26-119 -> (e_tot, e_1b, e_2b)."""
    W, M = phia.shape[0], phia.shape[1]
    omax = det_overlap_max(phia, phib, trials_up, trials_dn)
    Ga = np.zeros((W, M, M), dtype=np.complex128)
    Gb = np.zeros((W, M, M), dtype=np.complex128)
    e2b = np.zeros(W, dtype=np.complex128)
    ovlp = np.zeros(W, dtype=np.complex128)
    for d0, d1 in _chunks(trials_up.shape[0], det_chunk_size):
        Ghs_a, oas, Ghs_b, obs = [], [], [], []
        for d in range(d0, d1):
            gh, o = _gf_spin_single_det(phia, trials_up[d])
            Ghs_a.append(gh)
            oas.append(o)
            gh, o = _gf_spin_single_det(phib, trials_dn[d])
            Ghs_b.append(gh)
            obs.append(o)
        Gha, Ghb = np.stack(Ghs_a), np.stack(Ghs_b)
        oa, ob = np.stack(oas), np.stack(obs)
        Gha, Ghb = mask_small_overlap(Gha, Ghb, oa, ob, omax)
        w = ci_coeffs[d0:d1].conj()[:, None] * oa * ob
        ovlp = ovlp + np.sum(w, axis=0)
        w4 = w[:, :, None, None]
        Ga = Ga + np.sum(w4 * np.einsum('dpi,dwiq->dwpq', trials_up[d0:d1].conj(), Gha), axis=0)
        Gb = Gb + np.sum(w4 * np.einsum('dpi,dwiq->dwpq', trials_dn[d0:d1].conj(), Ghb), axis=0)
        e2 = []
        for k, d in enumerate(range(d0, d1)):
            ra = np.einsum('pi,gpq->giq', trials_up[d].conj(), chol)
            rb = np.einsum('pi,gpq->giq', trials_dn[d].conj(), chol)
            ec, ex = O.local_energy_2body(Gha[k], Ghb[k], ra, rb)
            e2.append(ec - ex)
        e2b = e2b + np.sum(w * np.stack(e2), axis=0)
    Ga = Ga / ovlp[:, None, None]
    Gb = Gb / ovlp[:, None, None]
    e_1b = O.local_energy_1body(h1e, Ga, Gb, enuc)
    e_2b = e2b / ovlp
    return e_1b + e_2b, e_1b, e_2b


def half_rotate_cholesky_multidet(chol, trials_up, trials_dn):
    """cholesky.py:470-510 -> rchols_a (ndet, G, Na, M), rchols_b."""
    ra = np.einsum('dpi,gpq->dgiq', trials_up.conj(), chol)
    rb = np.einsum('dpi,gpq->dgiq', trials_dn.conj(), chol)
    return ra, rb


def greens_function_multidet_force_bias(phia, phib, trials_up, trials_dn, ci_coeffs):
    """greens.py:314-369 -> (Ghalfa_all, Ghalfb_all, overlap, ovlp_a_all,
    ovlp_b_all): per-det Ghalf with non-finite entries set to 0, no
    small-overlap mask, one sum over all dets."""
    ga, oa = zip(*[_gf_spin_single_det(phia, t) for t in trials_up])
    gb, ob = zip(*[_gf_spin_single_det(phib, t) for t in trials_dn])
    Gha, Ghb = np.stack(ga), np.stack(gb)
    Gha = np.where(np.isfinite(Gha), Gha, 0.0)
    Ghb = np.where(np.isfinite(Ghb), Ghb, 0.0)
    oa, ob = np.stack(oa), np.stack(ob)
    ovlp = np.sum(ci_coeffs.conj()[:, None] * oa * ob, axis=0)
    return Gha, Ghb, ovlp, oa, ob


def local_energy_multidet_streamed(h1e_trials_a, h1e_trials_b, Ghalfa_all, Ghalfb_all, rchols_a,
                                   rchols_b, ci_coeffs, ovlp_a_all, ovlp_b_all, enuc,
                                   e_chunk_g=16):
    """energy.py:318-395 -> (e_tot, e_1b, e_2b): per-det one-body trace
    against the precontracted h1e_trials (ndet, M, N), per-det streamed
    two-body, both weighted by c_I^* o_Ia o_Ib over the summed overlap."""
    w = ci_coeffs.conj()[:, None] * ovlp_a_all * ovlp_b_all
    ovlp = np.sum(w, axis=0)
    e1 = (np.einsum('dpi,dwip->dw', h1e_trials_a, Ghalfa_all)
          + np.einsum('dpi,dwip->dw', h1e_trials_b, Ghalfb_all))
    e_1b = np.sum(w * e1, axis=0) / ovlp + enuc
    e2 = []
    for I in range(rchols_a.shape[0]):
        ec, ex = O.local_energy_2body_streamed(Ghalfa_all[I], Ghalfb_all[I], rchols_a[I],
                                               rchols_b[I], e_chunk_g)
        e2.append(ec - ex)
    e_2b = np.sum(w * np.stack(e2), axis=0) / ovlp
    return e_1b + e_2b, e_1b, e_2b


def build_propagator_multidet(h1e_mod, chol, trials_up, trials_dn, ci_coeffs, dt, rdm1=None):
    """afqmc_gto.py:1203-1225: propagator of the dominant det with the
    |c|^2-weighted (or given) charge density."""
    G = np.asarray(rdm1) if rdm1 is not None else g_charge(trials_up, trials_dn, ci_coeffs)
    return O.build_propagator(h1e_mod, chol, trials_up[0], trials_dn[0], dt, G_charge=G)


def propagate_walkers_multidet(phia, phib, weights, overlap, e_hybrid, propagator, chol,
                               trials_up, trials_dn, ci_coeffs, eshift, xi, fbbound=None,
                               exp_nmax=6, det_chunk_size=5):
    """afqmc_gto.py:728-839 with ``xi`` given -> (phia, phib, w, ovlp_new, e_hybrid)."""
    dt = propagator['dt']
    expH1 = propagator['expH1']
    mf_shift = propagator['mf_shift']
    if fbbound is None:
        fbbound = O.FBBOUND_DEFAULT
    Ga, Gb, ovlp = greens_function_multidet(phia, phib, trials_up, trials_dn, ci_coeffs,
                                            det_chunk_size)
    phia = np.matmul(expH1, phia)
    phib = np.matmul(expH1, phib)
    # einsum('gpq,wqp->gw', chol, G)
    G = chol.shape[0]
    W, M = Ga.shape[0], Ga.shape[1]
    Gt = (Ga + Gb).transpose(0, 2, 1).reshape(W, M * M)
    vbias = O._xr_matmul(Gt, chol.reshape(G, M * M).T).T
    xbar = (-np.sqrt(dt) * (1j * vbias - mf_shift[:, None])).T
    xabs = np.abs(xbar)
    with np.errstate(invalid='ignore', divide='ignore'):
        xbar = np.where(xabs > fbbound, xbar * fbbound / xabs, xbar)
    xshifted = xi - xbar
    cmf = -np.sqrt(dt) * np.einsum('wg,g->w', xshifted, mf_shift)
    cfb = np.sum(xi * xbar, axis=1) - 0.5 * np.sum(xbar * xbar, axis=1)
    VHS = O.build_vhs(xshifted, chol, dt)
    phia = O.apply_exp_vhs(VHS, phia, exp_nmax)
    phib = O.apply_exp_vhs(VHS, phib, exp_nmax)
    phia = np.matmul(expH1, phia)
    phib = np.matmul(expH1, phib)
    ovlp_new = greens_function_multidet_overlap_only(phia, phib, trials_up, trials_dn,
                                                     ci_coeffs, det_chunk_size)
    w_new, e_new = O.update_weights_phaseless(weights, ovlp, ovlp_new, cfb, cmf, e_hybrid,
                                              eshift, dt)
    return phia, phib, w_new, ovlp_new, e_new
