"""Memory-lean driver: counterpart of ``_AFQMCDriverGTO_EStream``
(OmegaQMC/afqmc_gto_estream.py:85-374).

Same set-up and the same step as :class:`~omegaqmc_amd.driver.AFQMCDriver`;
two things change, both for peak memory:

* the block-end energy is the streamed estimator -- single determinant:
  ``afq_local_energy_streamed`` with the one-body trace against the
  precontracted ``h1e @ trial^*`` (energy.py:218-252); multi-determinant:
  ``afq_greens_function_multidet_force_bias`` + ``afq_local_energy_multidet_streamed``
  (greens.py:314-369, energy.py:318-395) over per-det half-rotated vectors --
  the exchange in slabs of ``e_chunk_g`` vectors and no (W, M, M) Green's
  function;
* the step runs over walker chunks of ``walker_chunk_size`` walkers, so the
  step workspace (VHS and friends) is sized for one chunk.  The auxiliary
  fields of the whole population are drawn first from the step's Philox key
  and each chunk is propagated with its slice, so the trajectory is the
  unchunked one.  ``walker_chunk_size=None`` picks the largest power of two
  whose workspace fits ``mem_frac`` of the free device memory
  (afqmc_gto_estream.py:50-82); 0 forces the unchunked step.
"""

from __future__ import annotations

import sys
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .afqmc_gto import (C128, F64, PhiloxKey, WEIGHT_CLIP_FRACTION, aux_fields, do_binning_analysis,
                        greens_function_force_bias, greens_function_multidet_force_bias,
                        half_rotate_cholesky_multidet, local_energy_multidet_streamed,
                        local_energy_streamed)
from .driver import AFQMCDriver


def step_bytes_per_walker(nbasis, nup, ndown, naux, precision="fp64", ndet=0):
    """Device bytes of the step workspace per walker (leading terms of the
    library's layouts: VHS, propagator scratch, Green's function, fields,
    force-bias partials; multi-det adds the per-det inverses and G)."""
    M, Na, Nb, G = int(nbasis), int(nup), int(ndown), int(naux)
    N = max(Na, Nb)
    if precision == "fp64":
        b = 16 * M * M + 16 * M * (2 * Na + 2 * Nb) + 32 * M * N + 16 * M * (Na + Nb)
        b += G * (16 * 8 + 16 + 8)
    else:  # bf16x3 / fp32: complex VHS planes of 8 bytes per element
        b = 8 * M * M + 8 * M * (Na + Nb) * 2 + 16 * M * (Na + Nb) + G * (8 * 4 + 16 + 8)
    b += 16 * (Na * Na + Nb * Nb) + 64
    if ndet:
        b += 16 * ndet * (Na * Na + Nb * Nb) + 16 * M * (Na + Nb) + 32 * M * M + 48 * ndet
    return b


def autotune_walker_chunk_size(driver, free_bytes=None, mem_frac=0.25):
    """afqmc_gto_estream.py:50-82: largest power of two of walkers whose step
    workspace fits ``mem_frac`` of the free device memory, or None when the
    free memory is unknown."""
    if free_bytes is None:
        try:
            free_bytes, _ = torch.cuda.mem_get_info()
        except Exception:
            return None
    bpw = step_bytes_per_walker(driver.nbasis, driver.nup, driver.ndown, driver.naux,
                                driver.precision, driver.ci_coeffs.shape[0] if driver.multidet
                                else 0)
    n = max(1, int(free_bytes * mem_frac / bpw))
    p = 1
    while p * 2 <= n:
        p *= 2
    return p


class EStreamDriver(AFQMCDriver):
    """_AFQMCDriverGTO_EStream for array-input Hamiltonians (arguments of
    :class:`~omegaqmc_amd.driver.AFQMCDriver` plus ``e_chunk_g``)."""

    def __init__(self, integrals, e_chunk_g=16, **kw):
        super().__init__(integrals, **kw)
        self.e_chunk_g = int(e_chunk_g)
        h = self.h1e.to(C128)
        if self.multidet:
            # afqmc_gto_estream.py:122-136
            self.h1e_trials_a = torch.einsum('pq,dqi->dpi', h, self.trials_up.to(C128).conj())
            self.h1e_trials_b = torch.einsum('pq,dqi->dpi', h, self.trials_dn.to(C128).conj())
            self.rchols_a, self.rchols_b = half_rotate_cholesky_multidet(
                self.chol, self.trials_up, self.trials_dn)
        else:
            self.h1e_trial_a = (h @ self.trial_up.to(C128).conj()).contiguous()
            self.h1e_trial_b = (h @ self.trial_dn.to(C128).conj()).contiguous()
        self.walker_chunk_size = None

    def block_energy(self, st):
        """afqmc_gto_estream.py:321-341 (streamed, no full G); device values."""
        if self.multidet:
            Gha, Ghb, _, oa, ob = greens_function_multidet_force_bias(
                st['phia'], st['phib'], self.trials_up, self.trials_dn, self.ci_coeffs)
            e_tot, _, _ = local_energy_multidet_streamed(
                self.h1e_trials_a, self.h1e_trials_b, Gha, Ghb, self.rchols_a, self.rchols_b,
                self.ci_coeffs, oa, ob, self.enuc, self.e_chunk_g)
        else:
            Gha, Ghb, _ = greens_function_force_bias(st['phia'], st['phib'], self.trial_up,
                                                     self.trial_dn)
            e_tot, _, _ = local_energy_streamed(self.h1e_trial_a, self.h1e_trial_b, Gha, Ghb,
                                                self.rchol_a, self.rchol_b, self.enuc,
                                                self.e_chunk_g)
        W = st['phia'].shape[0]
        check(_lib.load().afq_block_stats(_lib.context().handle, W, ptr(st['weights']),
                                          ptr(e_tot), ptr(st['stats'])), "block_stats")
        return e_tot

    def _propagate(self, st, key, eshift):
        wc = self.walker_chunk_size
        W = st['phia'].shape[0]
        if not wc or wc >= W:
            return super()._propagate(st, key, eshift)
        # the whole population's fields from the step key, then chunk by chunk
        xi = aux_fields(key, W, self.naux)
        ov = torch.empty((W,), dtype=C128, device=st['phia'].device)
        for w0 in range(0, W, wc):
            w1 = min(w0 + wc, W)
            ov[w0:w1] = self.system.step(st['phia'][w0:w1], st['phib'][w0:w1],
                                         st['weights'][w0:w1], st['e_hybrid'][w0:w1], eshift,
                                         xi=xi[w0:w1])
        return ov

    def __call__(self, rng_key=None, num_walkers=100, num_blocks=100, num_steps_per_block=25,
                 stabilize_freq=5, pop_control_freq=5, num_blocks_equil=10, fname_log=None,
                 walker_chunk_size=None, **kw):
        if walker_chunk_size is None:
            wcs = autotune_walker_chunk_size(self)
            self.walker_chunk_size = wcs if (wcs is not None and wcs < num_walkers) else None
        elif walker_chunk_size == 0:
            self.walker_chunk_size = None
        else:
            self.walker_chunk_size = int(walker_chunk_size)
        if self.verbose and self.walker_chunk_size:
            print(f"  walker_chunk_size = {self.walker_chunk_size}", file=sys.stdout)
        out = super().__call__(rng_key=rng_key, num_walkers=num_walkers, num_blocks=num_blocks,
                               num_steps_per_block=num_steps_per_block,
                               stabilize_freq=stabilize_freq, pop_control_freq=pop_control_freq,
                               num_blocks_equil=num_blocks_equil, fname_log=fname_log, **kw)
        out['walker_chunk_size'] = self.walker_chunk_size
        return out


def get_afqmc_estream_func_from_integrals(h1, chol, ecore, nelec, e_chunk_g=16, **kw):
    """The EStream driver for MO-basis arrays (the reference builds it from a
    PySCF mean-field object with the arguments of get_afqmc_func)."""
    from .afqmc_gto import prepare_afqmc_integrals_from_arrays
    trial = kw.pop('trial', None)
    mo_coeff = kw.pop('mo_coeff', None)
    multidet_method = kw.pop('multidet_method', 'explicit')
    integrals = prepare_afqmc_integrals_from_arrays(h1, chol, ecore, nelec, mo_coeff=mo_coeff)
    return EStreamDriver(integrals, e_chunk_g=e_chunk_g, trial=trial,
                         multidet_method=None if trial is None else multidet_method, **kw)
