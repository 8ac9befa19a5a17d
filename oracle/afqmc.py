"""Single-determinant phaseless AFQMC, NumPy FP64 restatement (oracle).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function follows the reference function named in its docstring line by
line (complex128 everywhere, the same contraction and clipping order).  The
only deliberate difference: the reference draws the auxiliary fields ``xi``
and the comb offset ``r`` from ``jax.random`` inside the step; here they are
arguments, so the oracle and the HIP path can be fed identical numbers.

Pins (tests/test_oracle_afqmc.py):
  greens_*            S^-1 S = I, Tr G = N, G^2 = G, permutation-expansion det
                      (N <= 3), force-bias / overlap-only variants bit-equal
                      to the full variant (test_afqmc_kernels.py:97-134).
  local_energy*       phi = trial gives the closed-form determinant energy
                      (cholesky.py:448-467); brute-force 4-index ERI
                      contraction; L = 0 gives 2 sum eps_i + E_nuc; streamed
                      == unstreamed on the reference's own random harness
                      (test_afqmc_kernels.py:155-238).
  build_vhs/apply_exp_vhs   linearity in x, 6-term Taylor vs scipy expm within
                      the Taylor remainder bound, VHS = VHS^T.
  build_propagator    expm by eigendecomposition of the real symmetric
                      generator; mf_shift purely imaginary.
  update_weights_phaseless  dt -> 0 fixed point, clip and cos-projection edge
                      cases computed by hand.
  orthogonalize_walkers  Q^H Q = I, Q R = phi, R_jj > 0, det relation.
  population_control_comb  sum of new weights = W, sorted indices, copy
                      counts in {floor, ceil}, brute-force tooth search.
  force_bias_shift    vbias from full G (afqmc_gto.py:794-795) equals the
                      half-rotated form; cap and cfb/cmf by hand.
  do_binning_analysis uncorrelated/constant series closed forms.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

WEIGHT_CLIP_FRACTION = 0.10     # afqmc_gto.py:36
FBBOUND_DEFAULT = 1.0           # afqmc_gto.py:37


# ---------------------------------------------------------------------------
# Green's functions and overlaps (observables/greens.py)
# ---------------------------------------------------------------------------

def _slogdet_ovlp(S):
    sign, logdet = np.linalg.slogdet(S)
    return sign * np.exp(logdet), logdet


def greens_function_spin(phi, trial):
    """greens.py:12-46 -> (G, Ghalf, ovlp, log|ovlp|)."""
    S = np.einsum('wpi,pj->wij', phi, trial.conj())
    Sinv = np.linalg.inv(S)
    Ghalf = np.einsum('wij,wqj->wiq', Sinv, phi)
    G = np.einsum('pi,wiq->wpq', trial.conj(), Ghalf)
    ovlp, logdet = _slogdet_ovlp(S)
    return G, Ghalf, ovlp, logdet


def ghalf_ovlp_spin(phi, trial):
    """greens.py:49-76 -> (Ghalf, ovlp)."""
    S = np.einsum('wpi,pj->wij', phi, trial.conj())
    Sinv = np.linalg.inv(S)
    Ghalf = np.einsum('wij,wqj->wiq', Sinv, phi)
    ovlp, _ = _slogdet_ovlp(S)
    return Ghalf, ovlp


def overlap_only_spin(phi, trial):
    """greens.py:79-101."""
    S = np.einsum('wpi,pj->wij', phi, trial.conj())
    return _slogdet_ovlp(S)[0]


def greens_function_force_bias(phia, phib, trial_up, trial_dn):
    """greens.py:104-129 -> (Ghalfa, Ghalfb, overlap)."""
    Gha, oa = ghalf_ovlp_spin(phia, trial_up)
    Ghb, ob = ghalf_ovlp_spin(phib, trial_dn)
    return Gha, Ghb, oa * ob


def greens_function_overlap(phia, phib, trial_up, trial_dn):
    """greens.py:132-154."""
    return overlap_only_spin(phia, trial_up) * overlap_only_spin(phib, trial_dn)


def greens_function(phia, phib, trial_up, trial_dn):
    """greens.py:157-198 -> (Ga, Gb, Ghalfa, Ghalfb, overlap)."""
    Ga, Gha, oa, _ = greens_function_spin(phia, trial_up)
    Gb, Ghb, ob, _ = greens_function_spin(phib, trial_dn)
    return Ga, Gb, Gha, Ghb, oa * ob


# ---------------------------------------------------------------------------
# Local energy (observables/energy.py)
# ---------------------------------------------------------------------------

def local_energy_1body(h1e, Ga, Gb, enuc):
    """energy.py:32-48."""
    return (np.einsum('pq,wqp->w', h1e, Ga) + np.einsum('pq,wqp->w', h1e, Gb) + enuc)


def local_energy_2body(Ghalfa, Ghalfb, rchol_a, rchol_b):
    """energy.py:51-91 -> (e_coul, e_exch); exchange is sum_ij T_ij T_ji.

    X_g = einsum('giq,wiq->gw') and T = einsum('giq,wjq->gwij') are formed
    as tensordot (BLAS) products; the sums over (g, i, j) follow."""
    def coul(rc, gh):
        return np.tensordot(rc, gh, axes=([1, 2], [1, 2]))           # (G, W)

    def exch(rc, gh):
        T = np.tensordot(rc, gh, axes=([2], [2])).transpose(0, 2, 1, 3)  # (G, W, i, j)
        return 0.5 * np.sum(T * T.transpose(0, 1, 3, 2), axis=(0, 2, 3))

    e_coul = 0.5 * np.sum((coul(rchol_a, Ghalfa) + coul(rchol_b, Ghalfb)) ** 2, axis=0)
    e_exch = exch(rchol_a, Ghalfa) + exch(rchol_b, Ghalfb)
    return e_coul, e_exch


def local_energy(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc):
    """energy.py:94-130 -> (e_tot, e_1b, e_2b)."""
    e_1b = local_energy_1body(h1e, Ga, Gb, enuc)
    e_coul, e_exch = local_energy_2body(Ghalfa, Ghalfb, rchol_a, rchol_b)
    e_2b = e_coul - e_exch
    return e_1b + e_2b, e_1b, e_2b


def local_energy_2body_streamed(Ghalfa, Ghalfb, rchol_a, rchol_b, e_chunk_g=16):
    """energy.py:133-190 (slabbed over g)."""
    naux = rchol_a.shape[0]
    nw = Ghalfa.shape[0]
    e_coul = np.zeros(nw, dtype=np.result_type(Ghalfa, rchol_a))
    e_exch = np.zeros_like(e_coul)
    for g0 in range(0, naux, e_chunk_g):
        g1 = min(g0 + e_chunk_g, naux)
        ec, ex = local_energy_2body(Ghalfa, Ghalfb, rchol_a[g0:g1], rchol_b[g0:g1])
        e_coul = e_coul + ec
        e_exch = e_exch + ex
    return e_coul, e_exch


def local_energy_streamed(h1e_trial_a, h1e_trial_b, Ghalfa, Ghalfb, rchol_a, rchol_b,
                          enuc, e_chunk_g=16):
    """energy.py:193-252 (no full G: Tr(h G) = sum_pi (h Psi*)_pi Ghalf_ip)."""
    e_1b = (np.einsum('pi,wip->w', h1e_trial_a, Ghalfa)
            + np.einsum('pi,wip->w', h1e_trial_b, Ghalfb) + enuc)
    e_coul, e_exch = local_energy_2body_streamed(Ghalfa, Ghalfb, rchol_a, rchol_b, e_chunk_g)
    e_2b = e_coul - e_exch
    return e_1b + e_2b, e_1b, e_2b


# ---------------------------------------------------------------------------
# Setup (integrals/cholesky.py, afqmc_gto.py:574-627)
# ---------------------------------------------------------------------------

def half_rotate_cholesky(chol, trial_up, trial_dn):
    """cholesky.py:470-510: rchol[g,i,q] = sum_p trial[p,i]* L_g[p,q]."""
    return (np.einsum('pi,gpq->giq', trial_up.conj(), chol),
            np.einsum('pi,gpq->giq', trial_dn.conj(), chol))


def determinant_energy_from_integrals(h1, chol, ecore, nup, ndown):
    """cholesky.py:448-467 (closed-form energy of the aufbau determinant)."""
    e1 = np.trace(h1[:nup, :nup]) + np.trace(h1[:ndown, :ndown])
    ja = np.einsum('gii->g', chol[:, :nup, :nup])
    jb = np.einsum('gii->g', chol[:, :ndown, :ndown])
    ka = np.sum(chol[:, :nup, :nup] ** 2)
    kb = np.sum(chol[:, :ndown, :ndown] ** 2)
    e2 = 0.5 * (np.sum((ja + jb) ** 2) - ka - kb)
    return float(ecore + e1 + e2)


def build_propagator(h1e_mod, chol, trial_up, trial_dn, dt, G_charge=None):
    """afqmc_gto.py:574-627 -> {'expH1', 'mf_shift', 'dt'}."""
    if G_charge is None:
        G_charge = trial_up @ trial_up.T.conj() + trial_dn @ trial_dn.T.conj()
    mf_shift = 1j * np.einsum('gpq,qp->g', chol, G_charge)
    vhs_mf = 1j * np.einsum('g,gpq->pq', mf_shift, chol)
    H1 = h1e_mod - vhs_mf
    expH1 = scipy.linalg.expm(-0.5 * dt * H1)
    return {'expH1': expH1, 'mf_shift': mf_shift, 'dt': dt}


def e_const(enuc, mf_shift):
    """afqmc_gto.py:1245-1251."""
    return float(enuc + 0.5 * np.sum(mf_shift * mf_shift).real)


# ---------------------------------------------------------------------------
# Propagation (afqmc_gto.py:451-725)
# ---------------------------------------------------------------------------

def _xr_matmul(x, Lflat):
    """Complex (W,K) times real (K,P) as two real BLAS products."""
    if np.iscomplexobj(Lflat):
        return x @ Lflat
    return (x.real @ Lflat) + 1j * (x.imag @ Lflat)


def apply_exp_vhs(VHS, phi, nmax=6):
    """afqmc_gto.py:451-469: sum_{n<=nmax} VHS^n phi / n! (recursive terms).

    einsum('wpq,wqn->wpn') is the batched product VHS @ temp."""
    result = phi.copy()
    temp = phi.copy()
    for n in range(1, nmax + 1):
        temp = np.matmul(VHS, temp) / n
        result = result + temp
    return result


def build_vhs(xshifted, chol, dt, chol_chunk_g=None):
    """afqmc_gto.py:472-482: VHS = i sqrt(dt) sum_g xshifted[w,g] L_g.

    einsum('wg,gpq->wpq') is xshifted @ chol.reshape(G, M*M)."""
    naux, nb = chol.shape[0], chol.shape[1]
    step = naux if chol_chunk_g is None else chol_chunk_g
    W = xshifted.shape[0]
    VHS = np.zeros((W, nb, nb), dtype=np.complex128)
    sqdt = np.sqrt(dt)
    for g0 in range(0, naux, step):
        g1 = min(g0 + step, naux)
        part = _xr_matmul(xshifted[:, g0:g1], chol[g0:g1].reshape(g1 - g0, nb * nb))
        VHS = VHS + 1j * sqdt * part.reshape(W, nb, nb)
    return VHS


def update_weights_phaseless(weights, ovlp_old, ovlp_new, cfb, cmf, e_hybrid_old, eshift, dt):
    """afqmc_gto.py:527-571 -> (weights_new, e_hybrid_new)."""
    ovlp_ratio = ovlp_new / ovlp_old
    e_new = -(np.log(ovlp_ratio) + cfb + cmf) / dt
    ebound = np.sqrt(2.0 / dt)
    e_re = np.clip(e_new.real, eshift - ebound, eshift + ebound)
    e_new = e_re + 1j * e_new.imag
    avg = 0.5 * (e_new + e_hybrid_old)
    magn = np.abs(np.exp(-dt * (avg - eshift)))
    dtheta = (-dt * e_new - cfb).imag
    cosine = np.maximum(0.0, np.cos(dtheta))
    return weights * magn * cosine, e_new


def force_bias_shift(Ghalfa, Ghalfb, rchol_a, rchol_b, mf_shift, xi, dt, fbbound=None):
    """afqmc_gto.py:682-706 -> (xbar, xshifted, cfb, cmf, vbias)."""
    if fbbound is None:
        fbbound = FBBOUND_DEFAULT
    # einsum('giq,wiq->gw', rchol, Ghalf) = (Ghalf.reshape(W, N M) @ rchol.reshape(G, N M)^T)^T
    W, G = Ghalfa.shape[0], rchol_a.shape[0]
    vbias = np.zeros((W, G), dtype=np.complex128)
    for rc, gh in ((rchol_a, Ghalfa), (rchol_b, Ghalfb)):
        K = rc.shape[1] * rc.shape[2]
        if K:
            vbias = vbias + _xr_matmul(gh.reshape(W, K), rc.reshape(G, K).T)
    vbias = vbias.T
    xbar = (-np.sqrt(dt) * (1j * vbias - mf_shift[:, None])).T
    xabs = np.abs(xbar)
    with np.errstate(invalid='ignore', divide='ignore'):
        xbar = np.where(xabs > fbbound, xbar * fbbound / xabs, xbar)
    xshifted = xi - xbar
    cmf = -np.sqrt(dt) * np.einsum('wg,g->w', xshifted, mf_shift)
    cfb = np.sum(xi * xbar, axis=1) - 0.5 * np.sum(xbar * xbar, axis=1)
    return xbar, xshifted, cfb, cmf, vbias


def propagate_walkers(phia, phib, weights, overlap, e_hybrid, propagator, chol,
                      rchol_a, rchol_b, trial_up, trial_dn, eshift, xi,
                      fbbound=None, exp_nmax=6):
    """afqmc_gto.py:630-725 with the auxiliary fields ``xi`` (W, G) given.

    ``overlap`` is accepted and ignored, as in the reference (the old overlap
    is recomputed from the walkers at the start of the step).
    Returns (phia, phib, weights_new, overlap_new, e_hybrid_new).
    """
    dt = propagator['dt']
    expH1 = propagator['expH1']
    mf_shift = propagator['mf_shift']
    Gha, Ghb, ovlp = greens_function_force_bias(phia, phib, trial_up, trial_dn)
    # einsum('pq,wqn->wpn', expH1, phi) = expH1 @ phi (batched)
    phia = np.matmul(expH1, phia)
    phib = np.matmul(expH1, phib)
    _, xshifted, cfb, cmf, _ = force_bias_shift(Gha, Ghb, rchol_a, rchol_b, mf_shift,
                                                xi, dt, fbbound)
    VHS = build_vhs(xshifted, chol, dt)
    phia = apply_exp_vhs(VHS, phia, exp_nmax)
    phib = apply_exp_vhs(VHS, phib, exp_nmax)
    phia = np.matmul(expH1, phia)
    phib = np.matmul(expH1, phib)
    ovlp_new = greens_function_overlap(phia, phib, trial_up, trial_dn)
    w_new, e_new = update_weights_phaseless(weights, ovlp, ovlp_new, cfb, cmf,
                                            e_hybrid, eshift, dt)
    return phia, phib, w_new, ovlp_new, e_new


# ---------------------------------------------------------------------------
# Stabilisation and population control (afqmc_gto.py:126-238)
# ---------------------------------------------------------------------------

def _qr_batch(phi):
    """afqmc_gto.py:126-157: thin QR, columns times sign(diag R)."""
    Q, R = np.linalg.qr(phi)
    d = np.diagonal(R, axis1=1, axis2=2)
    with np.errstate(invalid='ignore', divide='ignore'):
        signs = np.where(d == 0, 0, d / np.abs(d))
    Qf = Q * signs[:, None, :]
    log_det = np.sum(np.log(np.abs(d)), axis=1)
    return Qf, log_det, np.prod(signs, axis=1)


def orthogonalize_walkers(phia, phib):
    """afqmc_gto.py:160-181 -> (phia, phib, log_detR, sign_detR)."""
    qa, la, sa = _qr_batch(phia)
    qb, lb, sb = _qr_batch(phib)
    return qa, qb, la + lb, sa * sb


def comb_indices(weights, r):
    """afqmc_gto.py:210-231 with a sequential FP64 running sum.

    ``total`` and the cumulative sum are accumulated left to right (the
    order the GPU kernel uses; XLA's scan order is unspecified).
    """
    w = np.asarray(weights, dtype=np.float64)
    nw = w.shape[0]
    total = np.cumsum(w)[-1]
    scale = total / nw
    ws = w / scale
    c = np.cumsum(ws)
    teeth = np.arange(nw, dtype=np.float64) + r
    idx = np.searchsorted(c, teeth, side='left')
    return np.clip(idx, 0, nw - 1)


def population_control_comb(weights, phia, phib, r, e_hybrid=None):
    """afqmc_gto.py:184-238 with the offset ``r`` given.

    Returns (weights_new, phia, phib, e_hybrid, idx).
    """
    idx = comb_indices(weights, r)
    wn = np.ones_like(np.asarray(weights, dtype=np.float64))
    eh = None if e_hybrid is None else e_hybrid[idx]
    return wn, phia[idx], phib[idx], eh, idx


def population_control_pair_branch(weights, phia, phib):
    """afqmc_gto.py:241-284 -> (weights_new, phia_new, phib_new, src).

    The reference orders the walkers with np.argsort's default (not stable)
    sort; here ties are ordered by index (kind='stable'), which agrees with it
    whenever the weights are distinct.  The total weight is summed left to
    right (the reference's jnp.sum may round the last bit differently, which
    matters only for a weight exactly at 0.1 <w> or 4 <w>).
    """
    w = np.array(weights, dtype=np.float64)
    pa, pb = np.array(phia), np.array(phib)
    n = w.shape[0]
    total = 0.0
    for x in w:
        total += x
    avg = total / n
    order = np.argsort(w, kind='stable')
    lo, hi = 0.1 * avg, 4.0 * avg
    src = np.arange(n, dtype=np.int32)
    for k in range(n // 2):
        il, ih = order[k], order[n - 1 - k]
        if w[ih] > hi and w[il] < lo:
            m = 0.5 * (w[il] + w[ih])
            w[il] = m
            w[ih] = m
            pa[il] = pa[ih]
            pb[il] = pb[ih]
            src[il] = ih
    return w, pa, pb, src


# ---------------------------------------------------------------------------
# Statistics (utils.py:86-149)
# ---------------------------------------------------------------------------

def do_binning_analysis(a):
    """utils.py:86-149 -> (mean, serr, std, kappa)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    xbar = a.mean()
    sdev = a.std()
    x = a - xbar
    ac = np.correlate(x, x, mode='full')
    ac = ac / ac[ac.shape[0] // 2]
    kappa = 1.0
    i = (ac.shape[0] >> 1) + 1
    while i < ac.shape[0] and ac[i] >= 0:
        kappa += 2.0 * ac[i]
        i += 1
    return xbar, sdev * (kappa / n) ** 0.5, sdev, kappa


# ---------------------------------------------------------------------------
# Driver loop (afqmc_gto.py:1257-1546), aux fields supplied by callables
# ---------------------------------------------------------------------------

class Driver:
    """Counterpart of ``_AFQMCDriverGTO`` for an already rotated system.

    ``xi_fn(step, nwalkers, naux)`` and ``r_fn(step)`` supply the auxiliary
    fields and comb offsets for 1-based step numbers.
    """

    def __init__(self, h1e, h1e_mod, chol, enuc, nup, ndown, dt=0.005, eshift0=None):
        self.h1e = np.asarray(h1e)
        self.h1e_mod = np.asarray(h1e_mod)
        self.chol = np.asarray(chol)
        self.enuc = float(enuc)
        self.nbasis = self.h1e.shape[0]
        self.nup, self.ndown = int(nup), int(ndown)
        self.dt = dt
        self.trial_up = np.eye(self.nbasis, self.nup)
        self.trial_dn = np.eye(self.nbasis, self.ndown)
        self.rchol_a, self.rchol_b = half_rotate_cholesky(self.chol, self.trial_up, self.trial_dn)
        self.propagator = build_propagator(self.h1e_mod, self.chol, self.trial_up,
                                           self.trial_dn, dt)
        self.e_const = e_const(self.enuc, self.propagator['mf_shift'])
        self.ehf = determinant_energy_from_integrals(self.h1e, self.chol, self.enuc,
                                                     self.nup, self.ndown)
        self.eshift0 = eshift0

    def block_energy(self, phia, phib, weights):
        Ga, Gb, Gha, Ghb, _ = greens_function(phia, phib, self.trial_up, self.trial_dn)
        e_tot, _, _ = local_energy(self.h1e, Ga, Gb, Gha, Ghb, self.rchol_a, self.rchol_b,
                                   self.enuc)
        w = np.abs(weights)
        return float(np.sum(w * e_tot.real) / np.sum(w)), e_tot

    def run(self, xi_fn, r_fn, num_walkers, num_blocks, num_steps_per_block,
            stabilize_freq=5, pop_control_freq=5, num_blocks_equil=0,
            record=None):
        W = num_walkers
        phia = np.tile(self.trial_up[None].astype(np.complex128), (W, 1, 1))
        phib = np.tile(self.trial_dn[None].astype(np.complex128), (W, 1, 1))
        weights = np.ones(W)
        overlap = np.ones(W, dtype=np.complex128)
        e_hybrid = np.zeros(W, dtype=np.complex128)
        eshift = 0.0 if self.eshift0 is None else float(self.eshift0) - self.e_const
        step = 0
        energies = []
        for _ in range(num_blocks_equil + num_blocks):
            acc_w = 0.0
            acc_e = 0.0
            for _ in range(num_steps_per_block):
                step += 1
                if step % stabilize_freq == 0:
                    phia, phib, log_detR, _ = orthogonalize_walkers(phia, phib)
                    overlap = overlap / np.exp(log_detR)
                xi = xi_fn(step, W, self.chol.shape[0])
                phia, phib, weights, overlap, e_hybrid = propagate_walkers(
                    phia, phib, weights, overlap, e_hybrid, self.propagator, self.chol,
                    self.rchol_a, self.rchol_b, self.trial_up, self.trial_dn, eshift, xi)
                if step > 1:
                    wbound = np.sum(np.abs(weights)) * WEIGHT_CLIP_FRACTION
                    weights = np.clip(weights, 0.0, wbound)
                if step % pop_control_freq == 0:
                    weights, phia, phib, e_hybrid, _ = population_control_comb(
                        weights, phia, phib, r_fn(step), e_hybrid)
                w_abs = np.abs(weights)
                acc_w += np.sum(w_abs)
                acc_e += np.sum(w_abs * e_hybrid.real)
                if record is not None:
                    record(step, phia, phib, weights, e_hybrid)
            e_block, _ = self.block_energy(phia, phib, weights)
            energies.append(e_block)
            if acc_w > 1e-10:
                eshift = acc_e / acc_w
        energies = np.array(energies)
        prod = energies[num_blocks_equil:]
        mean, err, std, kappa = do_binning_analysis(prod)
        return {'energy_blocks': energies, 'energy_mean': float(mean),
                'energy_err': float(err), 'energy_std': float(std),
                'kappa': float(kappa), 'ehf': self.ehf,
                'phia': phia, 'phib': phib, 'weights': weights, 'e_hybrid': e_hybrid}
