"""Plane-wave 3D electron gas (jellium) on the phaseless AFQMC step.

Counterpart of ``OmegaQMC/afqmc_pw_heg.py``: the system (k-grid shells,
box), the integrals (kinetic one-body term, symmetrised plane-wave Cholesky
vectors with their ERI sign, Madelung constant), the Fermi-sea trial and a
driver whose step and local energy are the library's FP64 step with the
vector signs set (``afq_system_set_chol_sign``): HS phase i for the symmetric
(sign +1) vectors and 1 for the antisymmetric (sign -1) ones, Coulomb and
exchange weighted by the sign.

The integral construction runs on the host in NumPy, as in the reference;
the orderings that decide the basis and vector order (k-points by |n|^2 from
an argsort of the same candidate list, q-vectors by |q|^2 over the set of
grid differences, (q, -q) pairs in that order) follow the reference so that
a given system has the same basis, trial and vector order.

The split-BF16 step needs symmetric L_g (its VHS is built on one triangle),
so systems with antisymmetric vectors run in FP64 or FP32-MFMA (its VHS
GEMM takes full M x M vectors; the block energy is the FP64 estimator);
``precision="bf16x3"`` raises for them.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr
from .afqmc_gto import C128, F64, System, _c, _r, _real_or_complex, half_rotate_cholesky
from .driver import AFQMCDriver


# ---------------------------------------------------------------------------
# system (afqmc_pw_heg.py:47-141)
# ---------------------------------------------------------------------------

def kgrid_shells(n_pw):
    """Integer k-points in complete shells of |n|^2 (afqmc_pw_heg.py:47-81).

    Returns (grid (N, 3) int, N): the shells are taken in increasing |n|^2
    while the count stays within ``n_pw`` (at least one shell).
    """
    nmax = int(np.ceil(n_pw ** (1.0 / 3.0))) + 5
    r = np.arange(-nmax, nmax + 1)
    cand = np.stack([a.ravel() for a in np.meshgrid(r, r, r, indexing='ij')], axis=1)
    n2 = np.sum(cand * cand, axis=1)
    order = np.argsort(n2)
    cand, n2 = cand[order], n2[order]
    norms, counts = np.unique(n2, return_counts=True)
    total = 0
    for c in counts:
        if total > 0 and total + c > n_pw:
            break
        total += int(c)
    return cand[:total], total


def build_3deg_system(rs, N_elec, N_pw, polarization='unpolarized', kappa=None):
    """Box and plane-wave basis of the electron gas (afqmc_pw_heg.py:84-140).

    V = N (4 pi / 3) rs^3, L = V^(1/3), k = (n + kappa) 2 pi / L.
    """
    volume = N_elec * (4.0 * np.pi / 3.0) * rs ** 3
    L = volume ** (1.0 / 3.0)
    grid, n_pw = kgrid_shells(N_pw)
    dk = 2.0 * np.pi / L
    if kappa is not None:
        kappa = np.asarray(kappa, dtype=np.float64)
        kvecs = (grid.astype(np.float64) + kappa[None, :]) * dk
    else:
        kvecs = grid.astype(np.float64) * dk
    if polarization == 'unpolarized':
        if N_elec % 2:
            raise ValueError("N_elec must be even for an unpolarized gas")
        nup = ndown = N_elec // 2
    elif polarization == 'polarized':
        nup, ndown = N_elec, 0
    else:
        raise ValueError(f"unknown polarization {polarization!r}")
    return {'rs': float(rs), 'N_elec': int(N_elec), 'N_pw': int(n_pw), 'nup': nup,
            'ndown': ndown, 'L': float(L), 'volume': float(volume), 'dk': float(dk),
            'kvecs': kvecs, 'grid_int': grid, 'kappa': kappa, 'polarization': polarization}


# ---------------------------------------------------------------------------
# integrals (afqmc_pw_heg.py:143-375)
# ---------------------------------------------------------------------------

def build_h1e_pw_3d(system):
    """Kinetic energy |k|^2 / 2 on the diagonal (afqmc_pw_heg.py:143-158)."""
    k = system['kvecs']
    return np.diag(0.5 * np.sum(k * k, axis=1))


def _qvectors(grid):
    """Nonzero differences G_i - G_j of the grid, ordered by |q|^2 over the
    set of differences (afqmc_pw_heg.py:199-212)."""
    diffs = set()
    n = grid.shape[0]
    g = [tuple(int(v) for v in row) for row in grid]
    for i in range(n):
        gi = g[i]
        for j in range(n):
            gj = g[j]
            q = (gi[0] - gj[0], gi[1] - gj[1], gi[2] - gj[2])
            if q != (0, 0, 0):
                diffs.add(q)
    return np.array(sorted(diffs, key=lambda t: t[0] * t[0] + t[1] * t[1] + t[2] * t[2]),
                    dtype=np.int64).reshape(-1, 3)


def build_cholesky_pw_3d(system):
    """Symmetrised Cholesky vectors of the 3D Coulomb kernel
    (afqmc_pw_heg.py:161-254).

    The density operator of momentum transfer q has L^q[k, k'] =
    sqrt(4 pi / (V |q|^2)) for G_k' = G_k - q.  Each (q, -q) pair, q first in
    the |q|^2 order, gives A = (L^q + L^-q) / sqrt 2 (symmetric, sign +1) and
    B = (L^q - L^-q) / sqrt 2 (antisymmetric, sign -1), so that
    (pq|rs) = sum_g sign_g L^g_pq L^g_rs.  A q equal to -q (not possible for
    nonzero integer q, kept for completeness) gives one A-type vector.

    Returns (chol (N_chol, N_pw, N_pw), q_vecs (N_chol, 3), chol_sign).
    """
    grid = np.asarray(system['grid_int'], dtype=np.int64)
    n_pw, volume, dk = system['N_pw'], system['volume'], system['dk']
    q_int = _qvectors(grid)
    nq = q_int.shape[0]
    if nq == 0:
        return (np.zeros((0, n_pw, n_pw)), np.zeros((0, 3)), np.zeros(0))
    # grid point -> basis index through a dense box lookup
    lo = grid.min(axis=0) - q_int.max(axis=0).clip(min=0) - 1
    span = (grid.max(axis=0) - q_int.min(axis=0).clip(max=0) + 1) - lo + 1
    box = -np.ones(tuple(span), dtype=np.int64)
    box[tuple((grid - lo).T)] = np.arange(n_pw)
    q_vecs = q_int.astype(np.float64) * dk
    pref = np.sqrt(4.0 * np.pi / (np.sum(q_vecs * q_vecs, axis=1) * volume))
    raw = np.zeros((nq, n_pw, n_pw))
    for iq in range(nq):
        tgt = grid - q_int[iq][None, :] - lo
        inside = np.all((tgt >= 0) & (tgt < span[None, :]), axis=1)
        jk = np.full(n_pw, -1, dtype=np.int64)
        jk[inside] = box[tuple(tgt[inside].T)]
        ok = jk >= 0
        raw[iq, np.nonzero(ok)[0], jk[ok]] = pref[iq]
    index = {tuple(int(v) for v in q): i for i, q in enumerate(q_int)}
    used = np.zeros(nq, dtype=bool)
    chol, sign, qv = [], [], []
    for iq in range(nq):
        if used[iq]:
            continue
        ineg = index.get(tuple(int(-v) for v in q_int[iq]))
        if ineg is not None and ineg != iq:
            chol.append((raw[iq] + raw[ineg]) / math.sqrt(2.0))
            sign.append(1.0)
            qv.append(q_vecs[iq])
            chol.append((raw[iq] - raw[ineg]) / math.sqrt(2.0))
            sign.append(-1.0)
            qv.append(q_vecs[iq])
            used[ineg] = True
        else:
            chol.append(raw[iq])
            sign.append(1.0)
            qv.append(q_vecs[iq])
        used[iq] = True
    return np.array(chol), np.array(qv), np.array(sign)


def compute_madelung_3d(system, nmax=10, eta=None):
    """Madelung energy per electron of the simple-cubic jellium cell by Ewald
    summation (afqmc_pw_heg.py:257-302):
    e_M = 1/2 [sum'_R erfc(eta R) / R + (4 pi / V) sum'_G exp(-G^2 / 4 eta^2) / G^2
               - 2 eta / sqrt(pi) - pi / (V eta^2)],  eta = sqrt(pi) / L by
    default (the sum does not depend on eta once both lattice sums converge)."""
    from scipy.special import erfc
    L, V = system['L'], system['volume']
    eta = math.sqrt(math.pi) / L if eta is None else float(eta)
    r = np.arange(-nmax, nmax + 1)
    n = np.stack([a.ravel() for a in np.meshgrid(r, r, r, indexing='ij')], axis=1)
    n2 = np.sum(n * n, axis=1)
    n2 = n2[n2 > 0].astype(np.float64)
    R = L * np.sqrt(n2)
    g2 = (2.0 * np.pi / L) ** 2 * n2
    e_real = np.sum(erfc(eta * R) / R)
    e_recip = np.sum((4.0 * np.pi / V) * np.exp(-g2 / (4.0 * eta * eta)) / g2)
    return 0.5 * (e_real + e_recip - 2.0 * eta / math.sqrt(math.pi) - math.pi / (V * eta * eta))


def prepare_3deg_integrals(system):
    """h1e, h1e_mod = h1e - 1/2 sum_g sign_g L_g L_g^T, the vectors, their
    signs and the Madelung constant (afqmc_pw_heg.py:305-342)."""
    h1e = build_h1e_pw_3d(system)
    chol, q_vecs, sign = build_cholesky_pw_3d(system)
    if chol.shape[0]:
        h1e_mod = h1e - 0.5 * np.einsum('g,gij,gkj->ik', sign, chol, chol)
    else:
        h1e_mod = h1e.copy()
    return {'h1e': h1e, 'h1e_mod': h1e_mod, 'chol': chol, 'chol_sign': sign, 'q_vecs': q_vecs,
            'e_madelung': float(compute_madelung_3d(system)), 'N_q': int(chol.shape[0])}


def build_trial_3deg(h1e, nup, ndown):
    """Fermi sea: the lowest-|k|^2 plane waves, column i occupying the i-th
    of them in the argsort of the diagonal (afqmc_pw_heg.py:345-375)."""
    order = np.argsort(np.diag(np.asarray(h1e)))
    M = h1e.shape[0]
    up = np.zeros((M, nup))
    up[order[:nup], np.arange(nup)] = 1.0
    dn = np.zeros((M, ndown))
    if ndown:
        dn[order[:ndown], np.arange(ndown)] = 1.0
    return up, dn


def hf_energy_3deg(h1e, chol, chol_sign, trial_up, trial_dn, enuc):
    """Energy of the trial determinant: sum_k occupied |k|^2 / 2 minus the
    sign-weighted exchange 1/2 sum_g sign_g Tr((L_g G_s)^2) per spin, plus the
    background (afqmc_pw_heg.py:602-628)."""
    e = enuc
    for t in (trial_up, trial_dn):
        if t.shape[1] == 0:
            continue
        G = t @ t.T
        e += float(np.trace(h1e @ G))
        if chol.shape[0]:
            T = np.einsum('gij,jk->gik', chol, G)
            e -= 0.5 * float(np.einsum('g,gij,gji->', chol_sign, T, T))
    return e


def local_energy_2body_pw(Ghalfa, Ghalfb, rchol_a, rchol_b, chol_sign):
    """afqmc_pw_heg.py:414-457 -> (e_coul, e_exch) with the ERI sign of each
    vector (afq_local_energy_2body_pw)."""
    Gha, Ghb = _c(Ghalfa), _c(Ghalfb)
    ra, dta = _real_or_complex(rchol_a)
    rb, dtb = _real_or_complex(rchol_b)
    if dta != dtb:
        ra, rb, dta = _c(ra), _c(rb), _lib.AFQ_COMPLEX
    sg = _r(chol_sign)
    W, Na, M = Gha.shape
    Nb, G = Ghb.shape[1], ra.shape[0]
    ec = torch.empty((W,), dtype=C128, device=Gha.device)
    ex = torch.empty((W,), dtype=C128, device=Gha.device)
    check(_lib.load().afq_local_energy_2body_pw(_lib.context().handle, W, M, Na, Nb, G, ptr(Gha),
                                                ptr(Ghb), ptr(ra), ptr(rb), dta, ptr(sg), ptr(ec),
                                                ptr(ex)), "local_energy_2body_pw")
    return ec, ex


def local_energy_pw(h1e, chol, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc, chol_sign):
    """afqmc_pw_heg.py:460-471 -> (e_tot, e_1b, e_2b) (``chol`` is not read,
    as in the reference)."""
    Ga, Gb, Gha, Ghb = _c(Ga), _c(Gb), _c(Ghalfa), _c(Ghalfb)
    h, dth = _real_or_complex(h1e)
    ra, dta = _real_or_complex(rchol_a)
    rb, dtb = _real_or_complex(rchol_b)
    if len({dth, dta, dtb}) > 1:
        h, ra, rb, dth = _c(h), _c(ra), _c(rb), _lib.AFQ_COMPLEX
    sg = _r(chol_sign)
    W, Na, M = Gha.shape
    Nb, G = Ghb.shape[1], ra.shape[0]
    out = [torch.empty((W,), dtype=C128, device=Gha.device) for _ in range(3)]
    check(_lib.load().afq_local_energy_pw(_lib.context().handle, W, M, Na, Nb, G, ptr(h), ptr(Ga),
                                          ptr(Gb), ptr(Gha), ptr(Ghb), ptr(ra), ptr(rb), dth,
                                          float(enuc), ptr(sg), ptr(out[0]), ptr(out[1]),
                                          ptr(out[2])), "local_energy_pw")
    return tuple(out)


# ---------------------------------------------------------------------------
# driver (afqmc_pw_heg.py:470-832)
# ---------------------------------------------------------------------------

def _expm_sym(A, dt):
    """expm(-dt/2 A) of a real symmetric A by eigendecomposition."""
    A = 0.5 * (A + A.T)
    ev, V = torch.linalg.eigh(A)
    return ((V * torch.exp(-0.5 * dt * ev)) @ V.T).contiguous()


class HEGDriver(AFQMCDriver):
    """Counterpart of ``_AFQMCDriverPWHEG`` (afqmc_pw_heg.py:470-832).

    The walkers, step, block energy and block loop are those of
    :class:`AFQMCDriver`; the system carries the vector signs.  A fully
    polarized gas (ndown = 0) has no beta determinant at all: the reference's
    one-column dummy beta walker is never propagated, its Cholesky rows and
    Green's function are zeroed and its overlap ratio is 1, which is what
    leaving beta out gives.
    """

    def __init__(self, system, dt=0.005, include_coulomb=True, verbose=True,
                 precision="fp64", fbbound=None, exp_nmax=6, device="cuda"):
        self.dt = float(dt)
        self.verbose = verbose
        self.precision = precision
        self.heg = system
        self.include_coulomb = include_coulomb
        ints = prepare_3deg_integrals(system)
        M, nup, ndown = system['N_pw'], system['nup'], system['ndown']
        self.nbasis, self.nup, self.ndown = M, nup, ndown
        self.e_madelung = ints['e_madelung']
        self.N_q = ints['N_q']
        coul = include_coulomb and self.N_q > 0
        self.enuc = self.e_madelung * system['N_elec'] if include_coulomb else 0.0
        dev = torch.device(device)
        t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=F64, device=dev)  # noqa: E731
        tu, td = build_trial_3deg(ints['h1e'], nup, ndown)
        self.h1e = t(ints['h1e'])
        self.h1e_mod = t(ints['h1e_mod'])
        chol = ints['chol'] if coul else np.zeros((0, M, M))
        sign = ints['chol_sign'] if coul else np.zeros(0)
        self.chol = t(chol)
        self.chol_sign = t(sign)
        self.naux = self.chol.shape[0]
        self.trial_up, self.trial_dn = t(tu), t(td)
        self.e_trial = hf_energy_3deg(ints['h1e'], chol, sign, tu, td, self.enuc)
        self.ehf = self.e_trial
        if coul:
            # mf_shift_g = phase_g Tr(L_g G), vhs_mf = sum_g phase_g mf_shift_g L_g
            # (afqmc_pw_heg.py:563-585); the phase is i for sign +1, 1 for -1
            Gc = (self.trial_up @ self.trial_up.T + self.trial_dn @ self.trial_dn.T)
            tr = torch.einsum('gpq,qp->g', self.chol, Gc)
            phase = torch.where(self.chol_sign > 0, torch.full_like(tr, 0.0) + 1j,
                                torch.ones_like(tr).to(C128))
            mf = phase * tr.to(C128)
            vmf = torch.einsum('g,gpq->pq', phase * mf, self.chol.to(C128))
            if float(torch.max(torch.abs(vmf.imag))) > 1e-12 * max(1.0, float(torch.max(torch.abs(vmf.real)))):
                raise ValueError("mean-field one-body term is not real")
            expH1 = _expm_sym(self.h1e_mod - vmf.real, self.dt)
        else:
            mf = torch.zeros(0, dtype=C128, device=dev)
            expH1 = _expm_sym(self.h1e, self.dt)
        self.propagator = {'expH1': expH1, 'mf_shift': mf.contiguous(), 'dt': self.dt}
        self.e_const = float(self.enuc + 0.5 * torch.sum(mf * mf).real) if coul else self.enuc
        self.eshift0 = None
        self.multidet = False
        if coul:
            self.rchol_a, self.rchol_b = half_rotate_cholesky(self.chol, self.trial_up,
                                                              self.trial_dn)
        else:
            self.rchol_a = torch.zeros((0, nup, M), dtype=F64, device=dev)
            self.rchol_b = torch.zeros((0, ndown, M), dtype=F64, device=dev)
        self.system = System(self.propagator, self.chol, self.rchol_a, self.rchol_b,
                             self.trial_up, self.trial_dn, fbbound=fbbound, exp_nmax=exp_nmax,
                             precision=precision, chol_sign=self.chol_sign if coul else None)
        if verbose:
            print(f"3DEG: N_pw={M} nup={nup} ndown={ndown} N_q={self.N_q} rs={system['rs']:.2f} "
                  f"E_trial={self.e_trial:.10f} E_madelung/elec={self.e_madelung:.10f}",
                  file=sys.stderr)

    def __call__(self, rng_key=None, num_walkers=100, num_blocks=100, num_steps_per_block=25,
                 stabilize_freq=5, pop_control_freq=5, num_blocks_equil=10,
                 fname_log=None, **kw):
        out = super().__call__(rng_key=rng_key, num_walkers=num_walkers, num_blocks=num_blocks,
                               num_steps_per_block=num_steps_per_block,
                               stabilize_freq=stabilize_freq, pop_control_freq=pop_control_freq,
                               num_blocks_equil=num_blocks_equil, fname_log=fname_log, **kw)
        n = self.heg['N_elec']
        out['e_trial'] = self.e_trial
        out['e_per_elec_ha'] = out['energy_mean'] / n
        out['e_per_elec_ry'] = 2.0 * out['energy_mean'] / n
        return out


def get_afqmc_3deg_func(system, dt=0.005, include_coulomb=True, verbose=True,
                        precision="fp64", fbbound=None, exp_nmax=6):
    """afqmc_pw_heg.py:948-967: a driver for the given 3DEG system."""
    return HEGDriver(system, dt=dt, include_coulomb=include_coulomb, verbose=verbose,
                     precision=precision, fbbound=fbbound, exp_nmax=exp_nmax)


# ---------------------------------------------------------------------------
# twist-averaged boundary conditions (afqmc_pw_heg.py:970-1094)
# ---------------------------------------------------------------------------

def _radical_inverse(i, base):
    r, f = 0.0, 1.0
    while i > 0:
        f /= base
        r += f * (i % base)
        i //= base
    return r


def generate_halton_twists_3d(n_twists):
    """Twists kappa in [-0.5, 0.5)^3: the unscrambled Halton points (radical
    inverses of 0, 1, 2, ... in bases 2, 3, 5) minus 1/2
    (afqmc_pw_heg.py:970-981)."""
    return np.array([[_radical_inverse(i, b) for b in (2, 3, 5)] for i in range(n_twists)],
                    dtype=np.float64).reshape(n_twists, 3) - 0.5


def twisted_system_3d(base_system, kappa):
    """Copy of a system with k = (n + kappa) dk (afqmc_pw_heg.py:984-1001)."""
    kappa = np.asarray(kappa, dtype=np.float64)
    tw = dict(base_system)
    tw['kvecs'] = (np.asarray(base_system['grid_int'], dtype=np.float64) + kappa[None, :]) \
        * base_system['dk']
    tw['kappa'] = kappa
    return tw


def run_twist_averaged_afqmc_3deg(base_system, n_twists=20, dt=0.005, num_walkers=100,
                                  num_blocks=50, num_steps_per_block=25, num_blocks_equil=10,
                                  verbose=True, precision="fp64"):
    """One AFQMC run per Halton twist, energies averaged over the twists and
    errors combined in quadrature / n_twists (afqmc_pw_heg.py:1004-1094)."""
    twists = generate_halton_twists_3d(n_twists)
    e, err, et = np.zeros(n_twists), np.zeros(n_twists), np.zeros(n_twists)
    for i, kap in enumerate(twists):
        d = HEGDriver(twisted_system_3d(base_system, kap), dt=dt, include_coulomb=True,
                      verbose=False, precision=precision)
        r = d(num_walkers=num_walkers, num_blocks=num_blocks,
              num_steps_per_block=num_steps_per_block, num_blocks_equil=num_blocks_equil)
        e[i], err[i], et[i] = r['energy_mean'], r['energy_err'], d.e_trial
        if verbose:
            print(f"twist {i}/{n_twists} kappa=({kap[0]:.4f}, {kap[1]:.4f}, {kap[2]:.4f}) "
                  f"E={e[i]:.8f} +/- {err[i]:.8f} E_HF={et[i]:.8f}", file=sys.stderr)
    n = base_system['N_elec']
    mean, emean = float(np.mean(e)), float(np.sqrt(np.sum(err ** 2)) / n_twists)
    etm = float(np.mean(et))
    return {'twists': twists, 'e_twists': e, 'e_err_twists': err, 'e_trial_twists': et,
            'energy_mean': mean, 'energy_err': emean, 'e_trial_mean': etm,
            'e_per_elec_ha': mean / n, 'e_per_elec_ry': 2.0 * mean / n,
            'e_trial_per_elec_ry': 2.0 * etm / n}


# ---------------------------------------------------------------------------
# infinite-gas reference energies (afqmc_pw_heg.py:1097-1169)
# ---------------------------------------------------------------------------

def _kf(rs, polarization):
    """Fermi wave vector: (9 pi / 4)^(1/3) / rs unpolarized, (9 pi / 2)^(1/3)
    / rs fully polarized."""
    return ((9.0 * np.pi / (4.0 if polarization == 'unpolarized' else 2.0)) ** (1.0 / 3.0)) / rs


def hf_kinetic_energy_per_elec_ry(rs, polarization='unpolarized'):
    """Kinetic energy per electron of the infinite Fermi sea, 3/5 kF^2 / 2 Ha,
    in Ry (afqmc_pw_heg.py:1097-1111)."""
    return 0.6 * _kf(rs, polarization) ** 2


def hf_exchange_energy_per_elec_ry(rs, polarization='unpolarized'):
    """Exchange energy per electron of the infinite Fermi sea, -3 kF / (4 pi)
    Ha, in Ry (afqmc_pw_heg.py:1114-1136)."""
    return 2.0 * (-3.0 / (4.0 * np.pi)) * _kf(rs, polarization)


# Perdew-Zunger (Phys. Rev. B 23, 5048 (1981), Table XII) fits of the
# Ceperley-Alder correlation energy, Hartree per electron: (gamma, beta1,
# beta2) for rs >= 1 and (A, B, C, D) for rs < 1.
_PZ = {'unpolarized': ((-0.1423, 1.0529, 0.3334), (0.0311, -0.048, 0.0020, -0.0116)),
       'polarized': ((-0.0843, 1.3981, 0.2611), (0.01555, -0.0269, 0.0007, -0.0048))}


def pz_correlation_energy(rs, polarization='unpolarized'):
    """Correlation energy per electron (Hartree) of the Perdew-Zunger fit
    (afqmc_pw_heg.py:1139-1169)."""
    (g, b1, b2), (a, b, c, d) = _PZ['unpolarized' if polarization == 'unpolarized'
                                   else 'polarized']
    if rs >= 1.0:
        return g / (1.0 + b1 * np.sqrt(rs) + b2 * rs)
    return a * np.log(rs) + b + c * rs * np.log(rs) + d * rs
