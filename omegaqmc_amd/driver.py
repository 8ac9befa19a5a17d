"""Device-resident AFQMC driver: counterpart of ``_AFQMCDriverGTO``
(OmegaQMC/afqmc_gto.py:1034-1631) for array-input Hamiltonians.

Per step (afqmc_gto.py:1383-1451): QR every ``stabilize_freq`` steps
(before the step), one fused phaseless step, weight clip (skipped at step
1), comb every ``pop_control_freq`` steps, on-device accumulators.  Per
block: mixed-estimator energy, Kish size, eshift update -- the only host
synchronisation.  Walker buffers are preallocated; steady-state steps
allocate nothing.
"""

from __future__ import annotations

import sys
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .afqmc_gto import (C128, F64, MultiDetSystem, PhiloxKey, System, WEIGHT_CLIP_FRACTION, _r,
                        build_propagator, determinant_energy_from_integrals, do_binning_analysis,
                        greens_function_multidet_overlap_only, greens_function_overlap,
                        half_rotate_cholesky, multidet_charge_density,
                        prepare_afqmc_integrals_from_arrays, walker_status)


class AFQMCDriver:
    """Set up once (integrals, trial, propagator), run many times."""

    def __init__(self, integrals, dt=0.005, e_hf=None, eshift0=None, verbose=True,
                 precision="fp64", fbbound=None, exp_nmax=6, trial=None, det_chunk_size=5,
                 multidet_method=None):
        self.dt = float(dt)
        self.verbose = verbose
        self.precision = precision
        self.h1e = integrals['h1e']
        self.h1e_mod = integrals['h1e_mod']
        self.chol = integrals['chol']
        self.enuc = float(integrals['enuc'])
        self.nbasis = int(integrals['nbasis'])
        self.nup = int(integrals['nup'])
        self.ndown = int(integrals['ndown'])
        self.naux = int(self.chol.shape[0])
        if e_hf is None:
            e_hf = determinant_energy_from_integrals(self.h1e, self.chol, self.enuc, self.nup,
                                                     self.ndown)
        self.ehf = float(e_hf)
        dev = self.chol.device
        self.multidet = trial is not None
        if self.multidet:
            # afqmc_gto.py:1163-1225: dets from occupation strings, the
            # dominant det initialises the walkers, the propagator uses the
            # |c|^2-weighted (or the given) charge density
            eye = torch.eye(self.nbasis, dtype=F64, device=dev)
            occ_up = torch.as_tensor(np.asarray(trial['occ_up']), dtype=torch.long, device=dev)
            occ_dn = torch.as_tensor(np.asarray(trial['occ_dn']), dtype=torch.long, device=dev)
            self.trials_up = eye[:, occ_up].permute(1, 0, 2).contiguous()
            self.trials_dn = eye[:, occ_dn].permute(1, 0, 2).contiguous()
            self.ci_coeffs = torch.as_tensor(np.asarray(trial['ci_coeffs']), device=dev).to(C128)
            self.trial_up, self.trial_dn = self.trials_up[0], self.trials_dn[0]
            if trial.get('rdm1') is not None:
                Gc = torch.as_tensor(np.asarray(trial['rdm1']), device=dev).to(C128)
            else:
                Gc = multidet_charge_density(self.trials_up, self.trials_dn, self.ci_coeffs)
            self.propagator = build_propagator(self.h1e_mod, self.chol, self.trial_up,
                                               self.trial_dn, self.dt, G_charge=Gc)
        else:
            self.trial_up = torch.eye(self.nbasis, self.nup, dtype=F64, device=dev)
            self.trial_dn = torch.eye(self.nbasis, self.ndown, dtype=F64, device=dev)
            self.rchol_a, self.rchol_b = half_rotate_cholesky(self.chol, self.trial_up,
                                                              self.trial_dn)
            self.propagator = build_propagator(self.h1e_mod, self.chol, self.trial_up,
                                               self.trial_dn, self.dt)
        mf = self.propagator['mf_shift']
        self.e_const = float(self.enuc + 0.5 * torch.sum(mf * mf).real)
        self.eshift0 = eshift0
        if self.multidet:
            self.system = MultiDetSystem(self.propagator, self.chol, self.trials_up,
                                         self.trials_dn, self.ci_coeffs, fbbound=fbbound,
                                         exp_nmax=exp_nmax, det_chunk_size=det_chunk_size,
                                         precision=precision, method=multidet_method)
        else:
            self.system = System(self.propagator, self.chol, self.rchol_a, self.rchol_b,
                                 self.trial_up, self.trial_dn, fbbound=fbbound,
                                 exp_nmax=exp_nmax, precision=precision)

    # -- buffers ------------------------------------------------------------
    def init_walkers(self, W):
        dev = self.chol.device
        M, Na, Nb = self.nbasis, self.nup, self.ndown
        st = {
            'phia': self.system.trial_up.unsqueeze(0).repeat(W, 1, 1).contiguous(),
            'phib': self.system.trial_dn.unsqueeze(0).repeat(W, 1, 1).contiguous(),
            'phia2': torch.empty((W, M, Na), dtype=C128, device=dev),
            'phib2': torch.empty((W, M, Nb), dtype=C128, device=dev),
            'weights': torch.ones(W, dtype=F64, device=dev),
            'weights2': torch.empty(W, dtype=F64, device=dev),
            'overlap': torch.ones(W, dtype=C128, device=dev),
            'e_hybrid': torch.zeros(W, dtype=C128, device=dev),
            'e_hybrid2': torch.empty(W, dtype=C128, device=dev),
            'log_detR': torch.empty(W, dtype=F64, device=dev),
            'idx': torch.empty(W, dtype=torch.int32, device=dev),
            'acc': torch.zeros(2, dtype=F64, device=dev),
            'stats': torch.zeros(4, dtype=F64, device=dev),
        }
        return st

    def step(self, st, step_count, seed, eshift, stabilize_freq, pop_control_freq,
             comb_idx=None):
        """One driver step (afqmc_gto.py:1387-1451) on preallocated buffers.

        ``comb_idx`` (device int32 (W,)) replays given survivor indices at a
        population-control step instead of drawing the comb -- used to follow
        another run's trajectory exactly (e.g. the FP64 run of the same fields).
        """
        L = _lib.load()
        ctx = _lib.context().handle
        W = st['phia'].shape[0]
        M, Na, Nb = self.nbasis, self.nup, self.ndown
        if step_count % stabilize_freq == 0:
            check(L.afq_orthogonalize_walkers(ctx, W, M, Na, Nb, ptr(st['phia']), ptr(st['phib']),
                                              ptr(st['log_detR']), None), "orthogonalize")
        key = PhiloxKey(seed, step_count)
        st['overlap'] = self._propagate(st, key, eshift)
        if step_count > 1:
            check(L.afq_clip_weights(ctx, W, ptr(st['weights']), WEIGHT_CLIP_FRACTION), "clip")
        if step_count % pop_control_freq == 0:
            if comb_idx is not None:
                st['idx'].copy_(comb_idx)
                check(L.afq_population_control_apply(ctx, W, M, Na, Nb, ptr(st['idx']),
                                                      ptr(st['phia']), ptr(st['phib']),
                                                      ptr(st['e_hybrid']), ptr(st['weights2']),
                                                      ptr(st['phia2']), ptr(st['phib2']),
                                                      ptr(st['e_hybrid2'])), "comb apply")
            else:
                r = float(L.afq_comb_offset(seed & (2 ** 64 - 1), step_count))
                check(L.afq_population_control_comb(ctx, W, M, Na, Nb, ptr(st['weights']),
                                                    ptr(st['phia']), ptr(st['phib']), r,
                                                    ptr(st['e_hybrid']), ptr(st['weights2']),
                                                    ptr(st['phia2']), ptr(st['phib2']),
                                                    ptr(st['e_hybrid2']), ptr(st['idx'])),
                      "comb")
            for a, b in (('phia', 'phia2'), ('phib', 'phib2'), ('weights', 'weights2'),
                         ('e_hybrid', 'e_hybrid2')):
                st[a], st[b] = st[b], st[a]
        check(L.afq_accumulate_hybrid(ctx, W, ptr(st['weights']), ptr(st['e_hybrid']),
                                      ptr(st['acc'])), "accumulate")

    def _propagate(self, st, key, eshift):
        """The fused step of every walker in place; returns overlap_new."""
        return self.system.step(st['phia'], st['phib'], st['weights'], st['e_hybrid'], eshift,
                                key=key)

    def block_energy(self, st):
        """Mixed-estimator block energy (afqmc_gto.py:1466-1496); device values."""
        e_tot, _, _ = self.system.local_energy(st['phia'], st['phib'], self.h1e, self.enuc)
        L = _lib.load()
        W = st['phia'].shape[0]
        check(L.afq_block_stats(_lib.context().handle, W, ptr(st['weights']), ptr(e_tot),
                                ptr(st['stats'])), "block_stats")
        return e_tot

    def walker_status(self, st, zero_overlap_tol=1e-300):
        """Per-walker flag words and summary (device) for the current
        walkers, with their overlaps recomputed (the state's overlap is the
        last step's, before any comb gather)."""
        if self.multidet:
            ov = greens_function_multidet_overlap_only(st['phia'], st['phib'], self.trials_up,
                                                       self.trials_dn, self.ci_coeffs)
        else:
            ov = greens_function_overlap(st['phia'], st['phib'], self.system.trial_up,
                                         self.system.trial_dn)
        return walker_status(st['phia'], st['phib'], ov, st['weights'], st['e_hybrid'],
                             zero_overlap_tol)

    def __call__(self, rng_key=None, num_walkers=100, num_blocks=100, num_steps_per_block=25,
                 stabilize_freq=5, pop_control_freq=5, num_blocks_equil=10, fname_log=None,
                 walker_chunk_size=None, eshift0=None, local_energy_cap=False,
                 check_status=False, graph=False):
        seed = rng_key.seed if isinstance(rng_key, PhiloxKey) else (42 if rng_key is None
                                                                   else int(rng_key))
        batch = None
        if graph:
            # the block's steps as hipGraph replays of a one-system afq_batch:
            # the same kernels, fields, comb offsets and energy shift as the
            # eager steps, so the blocks are bit-identical to graph=False
            if self.multidet or type(self)._propagate is not AFQMCDriver._propagate:
                raise ValueError("graph=True drives the single-determinant step only")
            from .batch import Batch
            batch = Batch([self.system], [num_walkers], seeds=[seed],
                          stabilize_freq=stabilize_freq, pop_control_freq=pop_control_freq,
                          nstreams=1)
        fout = sys.stdout if not fname_log else open(fname_log, 'w', 1)
        st = self.init_walkers(num_walkers)
        if eshift0 is None:
            eshift0 = self.eshift0
        eshift = 0.0 if eshift0 is None else float(eshift0) - self.e_const
        ebound = float(np.sqrt(2.0 / self.dt))
        step_count = 0
        energy_blocks = []
        status_blocks = []
        n_capped = 0
        t_prev = time.perf_counter()
        for iblock in range(num_blocks_equil + num_blocks):
            if batch is None:
                st['acc'].zero_()
                for _ in range(num_steps_per_block):
                    step_count += 1
                    self.step(st, step_count, seed, eshift, stabilize_freq, pop_control_freq)
            else:
                batch.zero_accumulators()
                batch.set_eshift([eshift])
                batch.run(num_steps_per_block, "graph")
                step_count += num_steps_per_block
                batch.copy_state_into(0, st)
            e_tot = self.block_energy(st)
            if local_energy_cap:
                e_ref = ((float(eshift0) if eshift0 is not None else self.ehf) if iblock == 0
                         else eshift + self.e_const)
                er = e_tot.real
                n_capped += int(torch.sum(torch.abs(er - e_ref) > ebound))
                er = torch.clamp(er, e_ref - ebound, e_ref + ebound)
                w = torch.abs(st['weights'])
                e_block = float(torch.sum(w * er) / torch.sum(w))
                stats = st['stats'].cpu().numpy()
            else:
                stats = st['stats'].cpu().numpy()
                e_block = float(stats[0])
            energy_blocks.append(e_block)
            if check_status:
                _, summ = self.walker_status(st)
                summ = summ.cpu().numpy()
                status_blocks.append((int(summ[0]) & 0xffffffff, int(summ[1])))
                if self.verbose and summ[1] > 0:
                    print(f"  block {iblock}: {int(summ[1])} walkers flagged "
                          f"(flags {int(summ[0]):#x})", file=fout)
            acc = st['acc'].cpu().numpy()
            if acc[0] > 1e-10:
                eshift = float(acc[1] / acc[0])
            if self.verbose:
                t_now = time.perf_counter()
                phase = "EQ" if iblock < num_blocks_equil else "  "
                kish = stats[1] * stats[1] / stats[2] if stats[2] > 0 else 0.0
                print(f"{phase}{iblock:4d} {e_block:16.10f} {eshift:16.10f} {stats[1]:12.4f} "
                      f"{stats[3]:12.4f} {kish:>16.8e} {t_now - t_prev:>16.6f}", file=fout)
                t_prev = t_now
        energy_blocks = np.array(energy_blocks)
        mean, err, std, kappa = do_binning_analysis(energy_blocks[num_blocks_equil:])
        if fout is not sys.stdout:
            fout.close()
        out = {'energy_blocks': energy_blocks, 'energy_mean': float(mean),
               'energy_err': float(err), 'energy_std': float(std), 'kappa': float(kappa),
               'ehf': self.ehf, 'state': st}
        if check_status:
            # per block: (OR of the walker flag words, walkers flagged)
            out['status_blocks'] = status_blocks
        if local_energy_cap:
            out['local_energy_cap_bound'] = ebound
            out['n_local_energy_capped'] = n_capped
        return out


def lockstep_energy_error(ref, low, num_walkers, num_steps, seed=42, every=10,
                          num_steps_per_block=25, stabilize_freq=5, pop_control_freq=10):
    """Block-energy error of driver ``low`` against driver ``ref`` (same
    system, typically FP64 and a reduced precision) on one trajectory.

    The protocol is the reference driver's (afqmc_gto.py:1383-1502): blocks
    of ``num_steps_per_block`` steps, eshift = the hybrid-energy average of
    the block (afqmc_gto.py:1497-1502) taken from ``ref``'s accumulators and
    used by both runs.  Both runs draw the same auxiliary fields (same Philox
    stream); at every comb ``low`` replays ``ref``'s survivor indices, so the
    populations stay walker-for-walker comparable and the difference is the
    reduced-precision propagation plus ``low``'s own energy estimator (a
    BF16x3 system evaluates its exchange and Coulomb in split BF16).  Every
    ``every`` steps and at every block end both mixed estimators are
    evaluated, uncapped and with the reference's local_energy_cap (real part
    clipped to E_ref +/- sqrt(2/dt), E_ref = E_HF in block 0, then eshift +
    e_const; afqmc_gto.py:1475-1485).

    Returns a dict: 'trace' [(step, |dE| raw, |dE| capped, n_outside,
    sigma)] in Ha, n_outside = walkers of ``ref`` beyond the cap, sigma = the
    statistical error of ``ref``'s block energy (the |w|-weighted standard
    deviation of the local energies over the square root of the Kish size
    (sum |w|)^2 / sum w^2); 'block_ends' the same at block ends; 'max_raw',
    'max_capped' over all checkpoints; 'max_block_raw', 'max_block_capped'
    over block ends; 'block_ends_signed' [(step, dE raw, dE capped)] with
    sign (low minus ref) and 'mean_block_raw', 'mean_block_capped' the error
    of the run's mean block energy, the estimate a driver run reports
    (afqmc_gto.py:1517-1525 averages the production block energies)."""
    W = int(num_walkers)
    sr, sl = ref.init_walkers(W), low.init_walkers(W)
    eshift = ref.ehf - ref.e_const
    ebound = float(np.sqrt(2.0 / ref.dt))
    trace, ends, signed = [], [], []
    iblock = 0
    for step in range(1, int(num_steps) + 1):
        if (step - 1) % num_steps_per_block == 0:
            sr['acc'].zero_()
            sl['acc'].zero_()
        ref.step(sr, step, seed, eshift, stabilize_freq, pop_control_freq)
        comb = sr['idx'] if step % pop_control_freq == 0 else None
        low.step(sl, step, seed, eshift, stabilize_freq, pop_control_freq, comb_idx=comb)
        end = step % num_steps_per_block == 0
        if step % every == 0 or end:
            e_ref = ref.ehf if iblock == 0 else eshift + ref.e_const
            vals = []
            nout, sigma = 0, 0.0
            for drv, st in ((ref, sr), (low, sl)):
                e = drv.block_energy(st).real
                w = torch.abs(st['weights'])
                raw = float(torch.sum(w * e) / torch.sum(w))
                ec = torch.clamp(e, e_ref - ebound, e_ref + ebound)
                cap = float(torch.sum(w * ec) / torch.sum(w))
                vals.append((raw, cap))
                if drv is ref:
                    nout = int(torch.sum(torch.abs(e - e_ref) > ebound))
                    var = float(torch.sum(w * (e - raw) ** 2) / torch.sum(w))
                    kish = float(torch.sum(w) ** 2 / torch.sum(w * w))
                    sigma = (var / kish) ** 0.5
            rec = (step, abs(vals[1][0] - vals[0][0]), abs(vals[1][1] - vals[0][1]), nout, sigma)
            trace.append(rec)
            if end:
                ends.append(rec)
                signed.append((step, vals[1][0] - vals[0][0], vals[1][1] - vals[0][1]))
        if end:
            acc = sr['acc'].cpu().numpy()
            if acc[0] > 1e-10:
                eshift = float(acc[1] / acc[0])
            iblock += 1
    out = {'trace': trace, 'block_ends': ends,
           'max_raw': max((r[1] for r in trace), default=0.0),
           'max_capped': max((r[2] for r in trace), default=0.0),
           'max_block_raw': max((r[1] for r in ends), default=0.0),
           'max_block_capped': max((r[2] for r in ends), default=0.0),
           'block_ends_signed': signed,
           'mean_block_raw': abs(float(np.mean([r[1] for r in signed]))) if signed else 0.0,
           'mean_block_capped': abs(float(np.mean([r[2] for r in signed]))) if signed else 0.0}
    return out


def get_afqmc_func_from_integrals(h1, chol, ecore, nelec, mo_coeff=None, e_hf=None, dt=0.005,
                                  trial=None, eshift0=None, verbose=True, chol_chunk_g=128,
                                  det_chunk_size=5, multidet_method='explicit',
                                  precision="fp64"):
    """afqmc_gto.py:1592-1631.  ``trial`` (multi-det) is a dict with 'ci_coeffs',
    'occ_up', 'occ_dn' (and optionally 'rdm1').  ``multidet_method``:
    'explicit' (one inverse per walker and det) or 'wick' (ratio updates
    relative to det 0: the step's overlaps and Green's function from one
    inverse per walker; needs every det within four orbitals of det 0)."""
    if trial is not None and multidet_method not in ('explicit', 'wick'):
        raise ValueError(f"multidet_method {multidet_method!r}: 'explicit' or 'wick'")
    if trial is not None and trial.get('mo_coeff') is not None:
        raise NotImplementedError("rotate the integrals into the trial basis (mo_coeff=None)")
    integrals = prepare_afqmc_integrals_from_arrays(h1, chol, ecore, nelec, mo_coeff=mo_coeff)
    return AFQMCDriver(integrals, dt=dt, e_hf=e_hf, eshift0=eshift0, verbose=verbose,
                       precision=precision, trial=trial, det_chunk_size=det_chunk_size,
                       multidet_method=None if trial is None else multidet_method)
