"""Cavity-QED phaseless AFQMC on MI355X: the ``OmegaQMC.qed_afqmc_gto`` operator API.

Walkers carry a photon coordinate q next to their determinants.  The step
(propagate_qed_walkers, qed_afqmc_gto.py:335-440) is the symmetric Trotter
product

    e^{-dt/2 H_ph} e^{-dt/2 H_1(q)} e^{-dt H_2} e^{-dt/2 H_1(q)} e^{-dt/2 H_ph}

in which H_2 is the electronic two-body step of :mod:`.afqmc_gto` on the
augmented Cholesky vectors (the dipole self-energy is one more vector,
integrals/qed.py:130-190) and H_1(q) = H1_0 + sqrt(Omega) q d splits into the
q-independent ``expH1_0`` and a per-walker Taylor series of the dipole
matrix.  Every piece runs in the HIP kernels of ``libomega_afqmc.so``:
``afq_qed_propagate_photon``, ``afq_apply_exp_scaled``,
``afq_qed_photon_weights``, ``afq_qed_local_energy`` and
``afq_gather_walker_values`` plus the electronic entry points.

Random draws: the reference splits a ``jax.random`` key three times per
step (photon, fields, photon).  Here one Philox block of the step's key,
(W, naux + 2) normals, supplies them: columns [0, naux) are the auxiliary
fields, column naux the first photon half step, column naux + 1 the second.
The draws can also be passed explicitly (``xi=``, ``eta=``).

The reference driver builds its integrals with PySCF (QED-HF orbitals,
``int1e_r`` / ``int1e_rr`` dipole and quadrupole integrals); here
:func:`prepare_qed_integrals_from_arrays` takes the MO-basis arrays.
"""

from __future__ import annotations

import sys
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .afqmc_gto import (C128, F64, FBBOUND_DEFAULT, WEIGHT_CLIP_FRACTION, PhiloxKey, _c, _ctx,
                        _L, _r, aux_fields, build_propagator, comb_offset, do_binning_analysis,
                        force_bias, greens_function, greens_function_force_bias,
                        greens_function_overlap, half_rotate_cholesky, _build_vhs,
                        _apply_exp_vhs, _update_weights_phaseless, apply_one_body,
                        population_control_comb, Walkers)

PHOTON_TAYLOR_ORDER = 4  # qed_afqmc_gto.py:387 and :424


# ---------------------------------------------------------------------------
# Integrals and propagator (integrals/qed.py, qed_afqmc_gto.py:204-259)
# ---------------------------------------------------------------------------

class QEDWalkers(Walkers):
    """qed_afqmc_gto.py:80-102: the walker container plus the photon
    displacement coordinate q of each walker (all q0)."""

    def __init__(self, trial_up, trial_dn, nwalkers, nbasis, nup, ndown, q0=0.0):
        super().__init__(trial_up, trial_dn, nwalkers, nbasis, nup, ndown)
        self.q = torch.full((self.nwalkers,), float(q0), dtype=F64, device=self.weights.device)


def prepare_qed_integrals_from_arrays(h1e, chol_mo, dip_mo, enuc, nelec, quad_mo=None,
                                      omega=None):
    """integrals/qed.py:95-205 for MO-basis arrays.

    ``dip_mo`` is lambda eps.r in the MO basis, ``quad_mo`` (optional) the
    lambda^2 (eps.r)^2 matrix of the proper dipole self-energy (added to h1e
    as 0.5 quad_mo).  Returns the reference's dict: h1e, h1e_mod_0 = h1e + v0
    with v0 = -1/2 A^T A over the augmented vectors, chol_qed = [chol; dip],
    dip_mo, enuc, nbasis, nup, ndown (and omega when given)."""
    h = _r(h1e)
    if quad_mo is not None:
        h = h + 0.5 * _r(quad_mo)
    L = _r(chol_mo)
    d = _r(dip_mo)
    M = h.shape[0]
    if L.shape[1:] != (M, M) or d.shape != (M, M):
        raise ValueError("chol_mo (G, M, M) and dip_mo (M, M) must match h1e (M, M)")
    chol_qed = torch.cat([L, d.unsqueeze(0)], dim=0).contiguous()
    A = chol_qed.transpose(1, 2).reshape(-1, M)
    v0 = -0.5 * (A.T @ A)
    nup, ndown = (int(nelec[0]), int(nelec[1])) if not np.isscalar(nelec) else (
        (int(nelec) + 1) // 2, int(nelec) // 2)
    out = {'h1e': h.contiguous(), 'h1e_mod_0': (h + v0).contiguous(), 'chol_qed': chol_qed,
           'dip_mo': d, 'enuc': float(enuc), 'nbasis': M, 'nup': nup, 'ndown': ndown}
    if omega is not None:
        out['omega'] = float(omega)
    return out


def build_qed_propagator(h1e_mod_0, chol_qed, trial_up, trial_dn, dt, chol_chunk_g=None,
                         G_charge=None):
    """qed_afqmc_gto.py:204-259 -> {'expH1_0', 'mf_shift', 'dt'}: the plain
    propagator's mean-field shift and exp(-dt/2 (h1e_mod_0 - vhs_mf)) over the
    augmented vectors (the q-dependent part is applied in the step)."""
    p = build_propagator(h1e_mod_0, chol_qed, trial_up, trial_dn, dt, G_charge=G_charge,
                         chol_chunk_g=chol_chunk_g)
    return {'expH1_0': p['expH1'], 'mf_shift': p['mf_shift'], 'dt': p['dt']}


def default_q0(dip_mo, omega, nup, ndown):
    """qed_afqmc_gto.py:729-734: q0 = -<D>/sqrt(Omega) of the identity trial
    (sum of the occupied diagonal dipole elements of both spins)."""
    dd = torch.diagonal(_r(dip_mo))
    d_exp = float(torch.sum(dd[:nup]) + torch.sum(dd[:ndown]))
    return -d_exp / float(np.sqrt(omega)) if omega > 0 else 0.0


# ---------------------------------------------------------------------------
# Photon pieces
# ---------------------------------------------------------------------------

def _photon_draws(key, W, naux):
    """(xi (W, naux), eta1 (W,), eta2 (W,)) from the step key's Philox block."""
    z = aux_fields(key, W, naux + 2)
    return z[:, :naux].contiguous(), z[:, naux].contiguous(), z[:, naux + 1].contiguous()


def propagate_photon(q, omega, dt, s, q0, rng_key=None, eta=None):
    """qed_afqmc_gto.py:262-311 -> (q_new, log_weight, rng_key).

    One half step (delta = omega dt / 2) of the harmonic-oscillator kernel,
    importance sampled with the Gaussian trial (s, q0).  The normals are
    ``eta`` (W,) when given, else column 0 of the key's Philox block (the key
    is then advanced)."""
    q = _r(q)
    W = q.shape[0]
    if eta is None:
        if rng_key is None:
            rng_key = PhiloxKey()
        rng_key, key = rng_key.split()
        eta = aux_fields(key, W, 1).reshape(W)
    eta = _r(eta)
    qn = torch.empty_like(q)
    lw = torch.empty_like(q)
    check(_L().afq_qed_propagate_photon(_ctx(), W, ptr(q), ptr(eta), float(omega), float(dt),
                                        float(s), float(q0), ptr(qn), ptr(lw)),
          "propagate_photon")
    return qn, lw, rng_key


def apply_exp_scaled(coeff, mat, phi, nmax=PHOTON_TAYLOR_ORDER):
    """sum_{n<=nmax} (coeff[w] mat)^n phi[w] / n! for real coeff (W,) and mat (M, M)."""
    cf, m, phi = _r(coeff), _r(mat), _c(phi)
    W, M, N = phi.shape
    out = torch.empty_like(phi)
    check(_L().afq_apply_exp_scaled(_ctx(), W, M, N, ptr(cf), ptr(m), ptr(phi), int(nmax),
                                    ptr(out)), "apply_exp_scaled")
    return out


def _apply_exp_vhs_photon(coeff, mat, phia, phib, nmax=PHOTON_TAYLOR_ORDER, chunk_size=None):
    """qed_afqmc_gto.py:314-332 -> (phia, phib) (walker chunking is a memory
    knob of the reference; one launch per spin here)."""
    return apply_exp_scaled(coeff, mat, phia, nmax), apply_exp_scaled(coeff, mat, phib, nmax)


def photon_weights(weights, log_w1, log_w2):
    """qed_afqmc_gto.py:436-438: weights * exp(clip(log_w1 + log_w2, -10, 10)) (new array)."""
    w = _r(weights).clone()
    a, b = _r(log_w1), _r(log_w2)
    check(_L().afq_qed_photon_weights(_ctx(), w.shape[0], ptr(w), ptr(a), ptr(b)),
          "photon_weights")
    return w


# ---------------------------------------------------------------------------
# Estimator, comb, step
# ---------------------------------------------------------------------------

def qed_local_energy(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc, q, dip_mo, omega, s,
                     q0):
    """qed_afqmc_gto.py:156-201 -> (e_tot, e_1b, e_2b, e_ph), each (W,) complex."""
    h, d = _r(h1e), _r(dip_mo)
    Ga, Gb, Gha, Ghb = _c(Ga), _c(Gb), _c(Ghalfa), _c(Ghalfb)
    ra, rb, q = _r(rchol_a), _r(rchol_b), _r(q)
    W, Na, M = Gha.shape
    Nb = Ghb.shape[1]
    G = ra.shape[0]
    out = [torch.empty((W,), dtype=C128, device=Gha.device) for _ in range(4)]
    check(_L().afq_qed_local_energy(_ctx(), W, M, Na, Nb, G, ptr(h), ptr(d), ptr(Ga), ptr(Gb),
                                    ptr(Gha), ptr(Ghb), ptr(ra), ptr(rb), float(enuc), ptr(q),
                                    float(omega), float(s), float(q0), ptr(out[0]), ptr(out[1]),
                                    ptr(out[2]), ptr(out[3])), "qed_local_energy")
    return tuple(out)


def gather_walker_values(idx, src):
    """dst[i] = src[idx[i]] on the device (the photon share of the comb)."""
    src = _r(src)
    idx = idx.to(device=src.device, dtype=torch.int32).contiguous()
    dst = torch.empty_like(src)
    check(_L().afq_gather_walker_values(_ctx(), src.shape[0], ptr(idx), ptr(src), ptr(dst)),
          "gather_walker_values")
    return dst


def population_control_comb_qed(weights, phia, phib, q, rng_key, e_hybrid=None,
                                return_indices=False):
    """qed_afqmc_gto.py:106-153 -> (weights_new, phia_new, phib_new, q_new[, e_hybrid_new]).

    ``rng_key`` is a PhiloxKey (its comb offset is drawn) or the offset r."""
    out = population_control_comb(weights, phia, phib, rng_key, e_hybrid=e_hybrid,
                                  return_indices=True)
    idx = out[-1]
    qn = gather_walker_values(idx, q)
    res = out[:3] + (qn,) + (out[3:-1] if e_hybrid is not None else ())
    return (res + (idx,)) if return_indices else res


def propagate_qed_walkers(phia, phib, q, weights, overlap, e_hybrid, propagator, chol_qed,
                          rchol_a, rchol_b, trial_up, trial_dn, eshift, rng_key, dip_mo, omega, s,
                          q0, fbbound=None, exp_nmax=6, walker_chunk_size=None,
                          chol_chunk_g=None, xi=None, eta=None):
    """qed_afqmc_gto.py:335-440
    -> (phia, phib, q, weights_new, overlap_new, e_hybrid_new, rng_key).

    ``overlap`` is ignored as in the reference (the old overlap is the one of
    the force-bias Green's function).  The draws come from the key's Philox
    block unless ``xi`` (W, naux) and ``eta`` (2, W) are given."""
    dt = float(propagator['dt'])
    E = propagator['expH1_0']
    if isinstance(E, torch.Tensor) and E.is_complex():
        if float(torch.max(torch.abs(E.imag))) != 0.0:
            raise ValueError("expH1_0 must be real (mf_shift purely imaginary)")
        E = E.real
    mf_shift = propagator['mf_shift']
    pa, pb = _c(phia), _c(phib)
    W = pa.shape[0]
    naux = int(chol_qed.shape[0])
    if rng_key is None:
        rng_key = PhiloxKey()
    if xi is None or eta is None:
        rng_key, key = rng_key.split()
        xi_k, e1, e2 = _photon_draws(key, W, naux)
        xi = xi_k if xi is None else xi
        if eta is None:
            eta = (e1, e2)
    eta1, eta2 = eta[0], eta[1]
    fb = FBBOUND_DEFAULT if fbbound is None else float(fbbound)

    Gha, Ghb, ovlp = greens_function_force_bias(pa, pb, trial_up, trial_dn)
    q, lw1, _ = propagate_photon(q, omega, dt, s, q0, eta=eta1)
    c1 = -0.5 * dt * float(np.sqrt(omega)) * q
    pa, pb = _apply_exp_vhs_photon(c1, dip_mo, pa, pb)
    pa, pb = apply_one_body(E, pa), apply_one_body(E, pb)
    _, xshifted, cfb, cmf = force_bias(Gha, Ghb, rchol_a, rchol_b, mf_shift, xi, dt, fb)
    VHS = _build_vhs(xshifted, chol_qed, dt)
    pa, pb = _apply_exp_vhs(VHS, pa, exp_nmax), _apply_exp_vhs(VHS, pb, exp_nmax)
    del VHS
    pa, pb = apply_one_body(E, pa), apply_one_body(E, pb)
    c2 = -0.5 * dt * float(np.sqrt(omega)) * q
    pa, pb = _apply_exp_vhs_photon(c2, dip_mo, pa, pb)
    q, lw2, _ = propagate_photon(q, omega, dt, s, q0, eta=eta2)
    ovlp_new = greens_function_overlap(pa, pb, trial_up, trial_dn)
    w_new, e_new = _update_weights_phaseless(weights, ovlp, ovlp_new, cfb, cmf, e_hybrid, eshift,
                                             dt)
    check(_L().afq_qed_photon_weights(_ctx(), W, ptr(w_new), ptr(lw1), ptr(lw2)),
          "photon_weights")
    return pa, pb, q, w_new, ovlp_new, e_new, rng_key


# ---------------------------------------------------------------------------
# Driver (qed_afqmc_gto.py:586-1061)
# ---------------------------------------------------------------------------

class QEDDriver:
    """_QEDAFQMCDriverGTO for array-input integrals (single-determinant trial
    of the lowest orbitals, the reference's non-multidet branch)."""

    def __init__(self, integrals, omega, dt=0.005, s=1.0, q0='auto', e_hf=None, verbose=True,
                 fbbound=None, exp_nmax=6):
        self.dt = float(dt)
        self.omega = float(omega)
        self.s = float(s)
        self.verbose = verbose
        self.fbbound = fbbound
        self.exp_nmax = int(exp_nmax)
        self.h1e = integrals['h1e']
        self.h1e_mod_0 = integrals['h1e_mod_0']
        self.chol_qed = integrals['chol_qed']
        self.dip_mo = integrals['dip_mo']
        self.enuc = float(integrals['enuc'])
        self.nbasis = int(integrals['nbasis'])
        self.nup = int(integrals['nup'])
        self.ndown = int(integrals['ndown'])
        self.naux = int(self.chol_qed.shape[0])
        self.ehf = None if e_hf is None else float(e_hf)
        if isinstance(q0, str) and q0 == 'auto':
            self.q0 = default_q0(self.dip_mo, self.omega, self.nup, self.ndown)
        else:
            self.q0 = float(q0)
        dev = self.chol_qed.device
        self.trial_up = torch.eye(self.nbasis, self.nup, dtype=F64, device=dev)
        self.trial_dn = torch.eye(self.nbasis, self.ndown, dtype=F64, device=dev)
        self.rchol_a, self.rchol_b = half_rotate_cholesky(self.chol_qed, self.trial_up,
                                                          self.trial_dn)
        self.propagator = build_qed_propagator(self.h1e_mod_0, self.chol_qed, self.trial_up,
                                               self.trial_dn, self.dt)

    def init_walkers(self, W):
        dev = self.chol_qed.device
        return {
            'phia': self.trial_up.to(C128).unsqueeze(0).repeat(W, 1, 1).contiguous(),
            'phib': self.trial_dn.to(C128).unsqueeze(0).repeat(W, 1, 1).contiguous(),
            'q': torch.full((W,), self.q0, dtype=F64, device=dev),
            'weights': torch.ones(W, dtype=F64, device=dev),
            'overlap': torch.ones(W, dtype=C128, device=dev),
            'e_hybrid': torch.zeros(W, dtype=C128, device=dev),
            'log_detR': torch.empty(W, dtype=F64, device=dev),
            'acc': torch.zeros(2, dtype=F64, device=dev),
            'stats': torch.zeros(4, dtype=F64, device=dev),
        }

    def step(self, st, step_count, seed, eshift, stabilize_freq, pop_control_freq):
        """One driver step (qed_afqmc_gto.py:891-949) in place on ``st``."""
        L = _L()
        ctx = _ctx()
        W = st['phia'].shape[0]
        M, Na, Nb = self.nbasis, self.nup, self.ndown
        if step_count % stabilize_freq == 0:
            check(L.afq_orthogonalize_walkers(ctx, W, M, Na, Nb, ptr(st['phia']), ptr(st['phib']),
                                              ptr(st['log_detR']), None), "orthogonalize")
        pa, pb, q, w, ov, eh, _ = propagate_qed_walkers(
            st['phia'], st['phib'], st['q'], st['weights'], st['overlap'], st['e_hybrid'],
            self.propagator, self.chol_qed, self.rchol_a, self.rchol_b, self.trial_up,
            self.trial_dn, eshift, PhiloxKey(seed, step_count), self.dip_mo, self.omega, self.s,
            self.q0, fbbound=self.fbbound, exp_nmax=self.exp_nmax)
        if step_count > 1:
            check(L.afq_clip_weights(ctx, W, ptr(w), WEIGHT_CLIP_FRACTION), "clip")
        if step_count % pop_control_freq == 0:
            r = comb_offset(PhiloxKey(seed, step_count))
            w, pa, pb, q, eh = population_control_comb_qed(w, pa, pb, q, r, e_hybrid=eh)
        st.update(phia=pa, phib=pb, q=q, weights=w, overlap=ov, e_hybrid=eh)
        check(L.afq_accumulate_hybrid(ctx, W, ptr(st['weights']), ptr(st['e_hybrid']),
                                      ptr(st['acc'])), "accumulate")

    def block_energy(self, st):
        """qed_afqmc_gto.py:952-983 -> (e_tot, e_1b, e_2b, e_ph) per walker (device)."""
        Ga, Gb, Gha, Ghb, _ = greens_function(st['phia'], st['phib'], self.trial_up,
                                              self.trial_dn)
        return qed_local_energy(self.h1e, Ga, Gb, Gha, Ghb, self.rchol_a, self.rchol_b,
                                self.enuc, st['q'], self.dip_mo, self.omega, self.s, self.q0)

    def __call__(self, rng_key=None, num_walkers=100, num_blocks=100, num_steps_per_block=25,
                 stabilize_freq=5, pop_control_freq=5, num_blocks_equil=10,
                 walker_chunk_size=None):
        seed = rng_key.seed if isinstance(rng_key, PhiloxKey) else (42 if rng_key is None
                                                                   else int(rng_key))
        st = self.init_walkers(int(num_walkers))
        eshift = 0.0
        step_count = 0
        energy_blocks, q_mean_blocks, e_ph_blocks = [], [], []
        t_prev = time.perf_counter()
        for iblock in range(num_blocks_equil + num_blocks):
            st['acc'].zero_()
            for _ in range(num_steps_per_block):
                step_count += 1
                self.step(st, step_count, seed, eshift, stabilize_freq, pop_control_freq)
            e_tot, _, _, e_ph = self.block_energy(st)
            w = torch.abs(st['weights'])
            w_sum = torch.sum(w)
            vals = torch.stack([torch.sum(w * e_tot.real) / w_sum,
                                torch.sum(w * st['q']) / w_sum,
                                torch.sum(w * e_ph.real) / w_sum,
                                w_sum, w_sum * w_sum / torch.sum(w * w),
                                st['acc'][0], st['acc'][1]]).cpu().numpy()
            e_block, q_mean, e_ph_block = float(vals[0]), float(vals[1]), float(vals[2])
            energy_blocks.append(e_block)
            q_mean_blocks.append(q_mean)
            e_ph_blocks.append(e_ph_block)
            if vals[5] > 1e-10:
                eshift = float(vals[6] / vals[5])
            if self.verbose:
                t_now = time.perf_counter()
                phase = "EQ" if iblock < num_blocks_equil else "  "
                print(f"{phase}{iblock:4d} {e_block:16.10f} {eshift:16.10f} {q_mean:10.4f} "
                      f"{e_ph_block:12.6f} {vals[3]:10.4f} {vals[4]:>16.8e} "
                      f"{t_now - t_prev:>16.6f}", file=sys.stdout)
                t_prev = t_now
        energy_blocks = np.array(energy_blocks)
        mean, err, std, kappa = do_binning_analysis(energy_blocks[num_blocks_equil:])
        q_prod = np.array(q_mean_blocks[num_blocks_equil:])
        e_ph_prod = np.array(e_ph_blocks[num_blocks_equil:])
        return {'energy_blocks': energy_blocks, 'energy_mean': float(mean),
                'energy_err': float(err), 'energy_std': float(std), 'kappa': float(kappa),
                'ehf': self.ehf, 'e_qed_hf': None, 'q0': float(self.q0),
                'q_mean': float(np.mean(q_prod)) if q_prod.size else float('nan'),
                'e_photon_mean': float(np.mean(e_ph_prod)) if e_ph_prod.size else float('nan'),
                'q_mean_blocks': np.array(q_mean_blocks), 'e_ph_blocks': np.array(e_ph_blocks),
                'state': st}


def get_qed_afqmc_func_from_integrals(h1e, chol_mo, dip_mo, enuc, nelec, omega, quad_mo=None,
                                      dt=0.005, s=1.0, q0='auto', e_hf=None, verbose=True):
    """qed_afqmc_gto.py:1063-1106 for MO-basis arrays (the reference's PySCF
    front end -- QED-HF orbitals, dipole and quadrupole integrals -- stays
    with the caller)."""
    integrals = prepare_qed_integrals_from_arrays(h1e, chol_mo, dip_mo, enuc, nelec,
                                                  quad_mo=quad_mo, omega=omega)
    return QEDDriver(integrals, omega, dt=dt, s=s, q0=q0, e_hf=e_hf, verbose=verbose)
