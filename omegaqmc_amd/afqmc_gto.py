"""Phaseless AFQMC on MI355X: the reference's ``OmegaQMC.afqmc_gto`` operator API.

Drop-in counterparts of the functions of ``OmegaQMC/afqmc_gto.py`` and
``OmegaQMC/observables/{greens,energy}.py``.  Arrays are torch tensors on a
ROCm device (NumPy inputs are copied to the current device); every
contraction runs in the HIP kernels of ``libomega_afqmc.so`` through the C
ABI -- there is no CPU or torch fallback.  Same names, argument order and
return tuples as the reference, with one documented difference: the
reference draws auxiliary fields inside the step from a ``jax.random`` key,
here ``rng_key`` is a :class:`PhiloxKey` (seed, step) of the counter-based
stream, or the fields are passed explicitly with ``xi=``.

Setup helpers (integral rotation, half rotation, expm of the one-body
generator) use torch linear algebra on the device; they run once per
driver, never inside a step.
"""

from __future__ import annotations

import dataclasses
import math
import sys
import time

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr

WEIGHT_CLIP_FRACTION = 0.10     # afqmc_gto.py:36
FBBOUND_DEFAULT = 1.0           # afqmc_gto.py:37

C128 = torch.complex128
F64 = torch.float64


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _c(x):
    """complex128, contiguous, on the current device."""
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    x = x.to(device=_dev(), dtype=C128) if (x.device.type != "cuda" or x.dtype != C128) else x
    return x.contiguous()


def _r(x):
    """float64, contiguous, on the current device."""
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    if x.is_complex():
        raise TypeError("expected a real array")
    x = x.to(device=_dev(), dtype=F64) if (x.device.type != "cuda" or x.dtype != F64) else x
    return x.contiguous()


def _real_or_complex(x):
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x))
    if x.is_complex():
        return _c(x), _lib.AFQ_COMPLEX
    return _r(x), _lib.AFQ_REAL


def _ctx():
    return _lib.context().handle


def _L():
    return _lib.load()


@dataclasses.dataclass(frozen=True)
class PhiloxKey:
    """Counter-based key: auxiliary fields of step ``step`` under ``seed``."""

    seed: int = 42
    step: int = 0

    def split(self):
        """(advanced key, key of the current step) -- jax.random.split's role."""
        return PhiloxKey(self.seed, self.step + 1), self


def aux_fields(key: PhiloxKey, nwalkers: int, naux: int):
    """xi (W, G) ~ N(0,1) of the Philox stream (seed, step) on the device."""
    xi = torch.empty((nwalkers, naux), dtype=F64, device=_dev())
    check(_L().afq_aux_fields(_ctx(), int(key.seed) & (2 ** 64 - 1), int(key.step), nwalkers, naux,
                              ptr(xi)), "afq_aux_fields")
    return xi


def comb_offset(key: PhiloxKey) -> float:
    return float(_L().afq_comb_offset(int(key.seed) & (2 ** 64 - 1), int(key.step)))


# ---------------------------------------------------------------------------
# Green's functions (observables/greens.py)
# ---------------------------------------------------------------------------

def greens_function_force_bias(phia, phib, trial_up, trial_dn):
    """greens.py:104-129 -> (Ghalfa, Ghalfb, overlap)."""
    phia, phib, tu, td = _c(phia), _c(phib), _c(trial_up), _c(trial_dn)
    W, M, Na = phia.shape
    Nb = phib.shape[2]
    Gha = torch.empty((W, Na, M), dtype=C128, device=phia.device)
    Ghb = torch.empty((W, Nb, M), dtype=C128, device=phia.device)
    ov = torch.empty((W,), dtype=C128, device=phia.device)
    check(_L().afq_greens_function_force_bias(_ctx(), W, M, Na, Nb, ptr(phia), ptr(phib), ptr(tu),
                                              ptr(td), ptr(Gha), ptr(Ghb), ptr(ov)),
          "greens_function_force_bias")
    return Gha, Ghb, ov


def greens_function_overlap(phia, phib, trial_up, trial_dn):
    """greens.py:132-154 -> overlap."""
    phia, phib, tu, td = _c(phia), _c(phib), _c(trial_up), _c(trial_dn)
    W, M, Na = phia.shape
    Nb = phib.shape[2]
    ov = torch.empty((W,), dtype=C128, device=phia.device)
    check(_L().afq_greens_function_overlap(_ctx(), W, M, Na, Nb, ptr(phia), ptr(phib), ptr(tu),
                                           ptr(td), ptr(ov)), "greens_function_overlap")
    return ov


def greens_function(phia, phib, trial_up, trial_dn):
    """greens.py:157-198 -> (Ga, Gb, Ghalfa, Ghalfb, overlap)."""
    phia, phib, tu, td = _c(phia), _c(phib), _c(trial_up), _c(trial_dn)
    W, M, Na = phia.shape
    Nb = phib.shape[2]
    dev = phia.device
    Ga = torch.empty((W, M, M), dtype=C128, device=dev)
    Gb = torch.empty((W, M, M), dtype=C128, device=dev)
    Gha = torch.empty((W, Na, M), dtype=C128, device=dev)
    Ghb = torch.empty((W, Nb, M), dtype=C128, device=dev)
    ov = torch.empty((W,), dtype=C128, device=dev)
    check(_L().afq_greens_function(_ctx(), W, M, Na, Nb, ptr(phia), ptr(phib), ptr(tu), ptr(td),
                                   ptr(Ga), ptr(Gb), ptr(Gha), ptr(Ghb), ptr(ov)),
          "greens_function")
    return Ga, Gb, Gha, Ghb, ov


# ---------------------------------------------------------------------------
# Local energy (observables/energy.py)
# ---------------------------------------------------------------------------

def local_energy_2body(Ghalfa, Ghalfb, rchol_a, rchol_b):
    """energy.py:51-91 -> (e_coul, e_exch)."""
    Gha, Ghb = _c(Ghalfa), _c(Ghalfb)
    ra, dt_a = _real_or_complex(rchol_a)
    rb, dt_b = _real_or_complex(rchol_b)
    if dt_a != dt_b:
        ra, rb, dt_a = _c(ra), _c(rb), _lib.AFQ_COMPLEX
    W, Na, M = Gha.shape
    Nb = Ghb.shape[1]
    G = ra.shape[0]
    ec = torch.empty((W,), dtype=C128, device=Gha.device)
    ex = torch.empty((W,), dtype=C128, device=Gha.device)
    check(_L().afq_local_energy_2body(_ctx(), W, M, Na, Nb, G, ptr(Gha), ptr(Ghb), ptr(ra), ptr(rb),
                                      dt_a, ptr(ec), ptr(ex)), "local_energy_2body")
    return ec, ex


def local_energy(h1e, Ga, Gb, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc):
    """energy.py:94-130 -> (e_tot, e_1b, e_2b)."""
    Ga, Gb, Gha, Ghb = _c(Ga), _c(Gb), _c(Ghalfa), _c(Ghalfb)
    h, dth = _real_or_complex(h1e)
    ra, dta = _real_or_complex(rchol_a)
    rb, dtb = _real_or_complex(rchol_b)
    if len({dth, dta, dtb}) > 1:
        h, ra, rb, dth = _c(h), _c(ra), _c(rb), _lib.AFQ_COMPLEX
    W, Na, M = Gha.shape
    Nb = Ghb.shape[1]
    G = ra.shape[0]
    out = [torch.empty((W,), dtype=C128, device=Gha.device) for _ in range(3)]
    check(_L().afq_local_energy(_ctx(), W, M, Na, Nb, G, ptr(h), ptr(Ga), ptr(Gb), ptr(Gha),
                                ptr(Ghb), ptr(ra), ptr(rb), dth, float(enuc), ptr(out[0]),
                                ptr(out[1]), ptr(out[2])), "local_energy")
    return tuple(out)


def local_energy_streamed(h1e_trial_a, h1e_trial_b, Ghalfa, Ghalfb, rchol_a, rchol_b, enuc,
                          e_chunk_g=16):
    """energy.py:218-252 -> (e_tot, e_1b, e_2b) without full G."""
    hta, htb = _c(h1e_trial_a), _c(h1e_trial_b)
    Gha, Ghb = _c(Ghalfa), _c(Ghalfb)
    ra, dta = _real_or_complex(rchol_a)
    rb, dtb = _real_or_complex(rchol_b)
    if dta != dtb:
        ra, rb, dta = _c(ra), _c(rb), _lib.AFQ_COMPLEX
    W, Na, M = Gha.shape
    Nb = Ghb.shape[1]
    G = ra.shape[0]
    out = [torch.empty((W,), dtype=C128, device=Gha.device) for _ in range(3)]
    check(_L().afq_local_energy_streamed(_ctx(), W, M, Na, Nb, G, ptr(hta), ptr(htb), ptr(Gha),
                                         ptr(Ghb), ptr(ra), ptr(rb), dta, float(enuc),
                                         int(e_chunk_g), ptr(out[0]), ptr(out[1]), ptr(out[2])),
          "local_energy_streamed")
    return tuple(out)


# ---------------------------------------------------------------------------
# Propagation pieces (afqmc_gto.py:451-725)
# ---------------------------------------------------------------------------

def _apply_exp_vhs(VHS, phi, nmax=6):
    """afqmc_gto.py:451-469 -> sum_{n<=nmax} VHS^n phi / n!."""
    VHS, phi = _c(VHS), _c(phi)
    W, M, N = phi.shape
    out = torch.empty_like(phi)
    check(_L().afq_apply_exp_vhs(_ctx(), W, M, N, ptr(VHS), ptr(phi), int(nmax), ptr(out)),
          "_apply_exp_vhs")
    return out


def _build_vhs(xshifted_w, chol, dt, chol_chunk_g=None):
    """afqmc_gto.py:472-482 -> VHS = i sqrt(dt) sum_g x[w,g] L_g (chunking is internal)."""
    xs, L = _c(xshifted_w), _r(chol)
    W, G = xs.shape
    M = L.shape[1]
    VHS = torch.empty((W, M, M), dtype=C128, device=xs.device)
    check(_L().afq_build_vhs(_ctx(), W, M, G, ptr(xs), ptr(L), float(dt), ptr(VHS)), "_build_vhs")
    return VHS


def _apply_exp_vhs_from_chol(xshifted, chol, phia, phib, dt, nmax=6, chunk_size=None,
                             chol_chunk_g=None):
    """afqmc_gto.py:485-524 (walker chunking is a memory knob; same result)."""
    VHS = _build_vhs(xshifted, chol, dt, chol_chunk_g)
    return _apply_exp_vhs(VHS, phia, nmax), _apply_exp_vhs(VHS, phib, nmax)


def _update_weights_phaseless(weights, ovlp_old, ovlp_new, cfb, cmf, e_hybrid_old, eshift, dt):
    """afqmc_gto.py:527-571 -> (weights_new, e_hybrid_new)."""
    w = _r(weights)
    oo, on, f, m, eh = (_c(v) for v in (ovlp_old, ovlp_new, cfb, cmf, e_hybrid_old))
    W = w.shape[0]
    wn = torch.empty_like(w)
    en = torch.empty_like(eh)
    check(_L().afq_update_weights_phaseless(_ctx(), W, ptr(w), ptr(oo), ptr(on), ptr(f), ptr(m),
                                            ptr(eh), float(eshift), float(dt), ptr(wn), ptr(en)),
          "_update_weights_phaseless")
    return wn, en


def force_bias(Ghalfa, Ghalfb, rchol_a, rchol_b, mf_shift, xi, dt, fbbound=None):
    """afqmc_gto.py:682-706 -> (xbar, xshifted, cfb, cmf)."""
    Gha, Ghb = _c(Ghalfa), _c(Ghalfb)
    ra, rb, mf, xi = _r(rchol_a), _r(rchol_b), _c(mf_shift), _r(xi)
    W, Na, M = Gha.shape
    Nb = Ghb.shape[1]
    G = ra.shape[0]
    fb = FBBOUND_DEFAULT if fbbound is None else float(fbbound)
    dev = Gha.device
    xbar = torch.empty((W, G), dtype=C128, device=dev)
    xs = torch.empty((W, G), dtype=C128, device=dev)
    cfb = torch.empty((W,), dtype=C128, device=dev)
    cmf = torch.empty((W,), dtype=C128, device=dev)
    check(_L().afq_force_bias(_ctx(), W, M, Na, Nb, G, ptr(ra), ptr(rb), ptr(Gha), ptr(Ghb),
                              ptr(mf), ptr(xi), float(dt), fb, ptr(xbar), ptr(xs), ptr(cfb),
                              ptr(cmf)), "force_bias")
    return xbar, xs, cfb, cmf


def apply_one_body(expH1, phi):
    """afqmc_gto.py:678-679: einsum('pq,wqn->wpn', expH1, phi) with real expH1."""
    E, phi = _r(expH1), _c(phi)
    W, M, N = phi.shape
    out = torch.empty_like(phi)
    check(_L().afq_apply_one_body(_ctx(), W, M, N, ptr(E), ptr(phi), ptr(out)), "apply_one_body")
    return out


class Walkers:
    """afqmc_gto.py:76-123: nwalkers copies of the trial (complex128 device
    tensors), weights 1, overlaps 1, hybrid energies 0, log_detR and
    log_shift 0."""

    def __init__(self, trial_up, trial_dn, nwalkers, nbasis, nup, ndown):
        tu, td = _c(trial_up), _c(trial_dn)
        if tu.shape != (nbasis, nup) or td.shape != (nbasis, ndown):
            raise ValueError("trial shapes do not match (nbasis, nup) / (nbasis, ndown)")
        self.nwalkers, self.nbasis, self.nup, self.ndown = int(nwalkers), nbasis, nup, ndown
        dev = tu.device
        self.phia = tu.unsqueeze(0).repeat(self.nwalkers, 1, 1).contiguous()
        self.phib = td.unsqueeze(0).repeat(self.nwalkers, 1, 1).contiguous()
        self.weights = torch.ones(self.nwalkers, dtype=F64, device=dev)
        self.overlap = torch.ones(self.nwalkers, dtype=C128, device=dev)
        self.e_hybrid = torch.zeros(self.nwalkers, dtype=C128, device=dev)
        self.log_detR = torch.zeros(self.nwalkers, dtype=F64, device=dev)
        self.log_shift = torch.zeros(self.nwalkers, dtype=F64, device=dev)


def orthogonalize_walkers(phia, phib):
    """afqmc_gto.py:160-181 -> (phia, phib, log_detR, sign_detR) (new arrays)."""
    pa, pb = _c(phia).clone(), _c(phib).clone()
    W, M, Na = pa.shape
    Nb = pb.shape[2]
    ld = torch.empty((W,), dtype=F64, device=pa.device)
    sg = torch.empty((W,), dtype=C128, device=pa.device)
    check(_L().afq_orthogonalize_walkers(_ctx(), W, M, Na, Nb, ptr(pa), ptr(pb), ptr(ld), ptr(sg)),
          "orthogonalize_walkers")
    return pa, pb, ld, sg


def population_control_comb(weights, phia, phib, rng_key, e_hybrid=None, return_indices=False):
    """afqmc_gto.py:184-238 -> (weights_new, phia_new, phib_new[, e_hybrid_new]).

    ``rng_key`` is a PhiloxKey (its comb offset is drawn) or the offset r itself.
    """
    w, pa, pb = _r(weights), _c(phia), _c(phib)
    W, M, Na = pa.shape
    Nb = pb.shape[2]
    r = comb_offset(rng_key) if isinstance(rng_key, PhiloxKey) else float(rng_key)
    eh = _c(e_hybrid) if e_hybrid is not None else None
    wn = torch.empty_like(w)
    pan = torch.empty_like(pa)
    pbn = torch.empty_like(pb)
    ehn = torch.empty_like(eh) if eh is not None else None
    idx = torch.empty((W,), dtype=torch.int32, device=pa.device)
    check(_L().afq_population_control_comb(_ctx(), W, M, Na, Nb, ptr(w), ptr(pa), ptr(pb), r,
                                           ptr(eh), ptr(wn), ptr(pan), ptr(pbn), ptr(ehn), ptr(idx)),
          "population_control_comb")
    out = (wn, pan, pbn) if eh is None else (wn, pan, pbn, ehn)
    return (out + (idx,)) if return_indices else out


def population_control_pair_branch(weights, phia, phib, rng_key=None, return_sources=False):
    """afqmc_gto.py:241-284 -> (weights_new, phia_new, phib_new[, src]).

    The lightest and heaviest walkers are paired; a heavy walker above 4<w>
    paired with a light one below 0.1<w> replaces it, both at the mean
    weight.  ``rng_key`` is unused, as in the reference.  Ties in weight are
    ordered by walker index (a stable argsort)."""
    w, pa, pb = _r(weights).clone(), _c(phia).clone(), _c(phib).clone()
    W, M, Na = pa.shape
    Nb = pb.shape[2]
    src = torch.empty((W,), dtype=torch.int32, device=pa.device)
    check(_L().afq_population_control_pair_branch(_ctx(), W, M, Na, Nb, ptr(w), ptr(pa), ptr(pb),
                                                  ptr(src)), "population_control_pair_branch")
    return (w, pa, pb, src) if return_sources else (w, pa, pb)


# ---------------------------------------------------------------------------
# Setup (integrals/cholesky.py, afqmc_gto.py:574-627)
# ---------------------------------------------------------------------------

def one_body_chol_correction(chol):
    """cholesky.py:261-275: v0 = -1/2 A^T A, A = L.transpose(0,2,1).reshape(G M, M)."""
    L = _r(chol)
    A = L.transpose(1, 2).reshape(-1, L.shape[1])
    return -0.5 * (A.T @ A)


def prepare_afqmc_integrals_from_arrays(h1, chol, ecore, nelec, mo_coeff=None):
    """cholesky.py:392-445 (device arrays)."""
    h1 = _r(h1)
    L = _r(chol)
    norb = h1.shape[0]
    if h1.shape != (norb, norb) or L.ndim != 3 or tuple(L.shape[1:]) != (norb, norb):
        raise ValueError(f"shape mismatch: h1 {tuple(h1.shape)}, chol {tuple(L.shape)}")
    if not torch.allclose(h1, h1.T, atol=1e-10):
        raise ValueError("h1 is not symmetric")
    if not torch.allclose(L, L.transpose(1, 2), atol=1e-10):
        raise ValueError("Cholesky vectors are not symmetric in (p, q)")
    if mo_coeff is None:
        C = torch.eye(norb, dtype=F64, device=h1.device)
    else:
        C = _r(mo_coeff)
        h1 = C.T @ h1 @ C
        L = torch.matmul(torch.matmul(C.T.unsqueeze(0), L), C.unsqueeze(0)).contiguous()
    v0 = one_body_chol_correction(L)
    return {'h1e': h1, 'h1e_mod': h1 + v0, 'chol': L, 'enuc': float(ecore), 'nbasis': norb,
            'nup': int(nelec[0]), 'ndown': int(nelec[1]), 'mo_coeff': C}


def determinant_energy_from_integrals(h1, chol, ecore, nup, ndown):
    """cholesky.py:448-467."""
    h1, L = _r(h1), _r(chol)
    e1 = torch.trace(h1[:nup, :nup]) + torch.trace(h1[:ndown, :ndown])
    ja = torch.einsum('gii->g', L[:, :nup, :nup])
    jb = torch.einsum('gii->g', L[:, :ndown, :ndown])
    ka = torch.sum(L[:, :nup, :nup] ** 2)
    kb = torch.sum(L[:, :ndown, :ndown] ** 2)
    e2 = 0.5 * (torch.sum((ja + jb) ** 2) - ka - kb)
    return float(ecore + e1 + e2)


def half_rotate_cholesky(chol, trial_up, trial_dn, chunk_g=None):
    """cholesky.py:470-510 -> rchol_a (G,Na,M), rchol_b (G,Nb,M).

    Real trials give real rchol (the driver path); complex trials complex.
    """
    L = _r(chol)
    out = []
    for tr in (trial_up, trial_dn):
        if isinstance(tr, np.ndarray):
            tr = torch.from_numpy(tr)
        tr = tr.to(_dev())
        if tr.is_complex():
            r = torch.matmul(tr.conj().T.unsqueeze(0).to(C128), L.to(C128))
        else:
            r = torch.matmul(tr.to(F64).T.unsqueeze(0), L)
        out.append(r.contiguous())
    return out[0], out[1]


def build_propagator(h1e_mod, chol, trial_up, trial_dn, dt, G_charge=None, chol_chunk_g=None):
    """afqmc_gto.py:574-627 -> {'expH1' (real), 'mf_shift' (complex, imaginary), 'dt'}."""
    L = _r(chol)
    h = _r(h1e_mod)
    if G_charge is None:
        tu, td = _c(trial_up), _c(trial_dn)
        Gc = tu @ tu.conj().T + td @ td.conj().T
    else:
        Gc = _c(G_charge)
    G, M = L.shape[0], L.shape[1]
    # mf_shift[g] = i sum_pq L[g,p,q] Gc[q,p]  (Gc Hermitian, L real symmetric)
    Lf = L.reshape(G, M * M)
    tr = Lf.to(C128) @ Gc.T.reshape(M * M)
    mf_shift = 1j * tr
    vhs_mf = (1j * mf_shift) @ Lf.to(C128)
    H1 = h.to(C128) - vhs_mf.reshape(M, M)
    if torch.max(torch.abs(H1.imag)) > 1e-12 * max(1.0, float(torch.max(torch.abs(H1.real)))):
        expH1 = torch.linalg.matrix_exp(-0.5 * dt * H1)
    else:
        Hr = H1.real.contiguous()
        Hr = 0.5 * (Hr + Hr.T)
        ev, V = torch.linalg.eigh(Hr)
        expH1 = (V * torch.exp(-0.5 * dt * ev)) @ V.T
    return {'expH1': expH1.contiguous(), 'mf_shift': mf_shift.contiguous(), 'dt': float(dt)}


# ---------------------------------------------------------------------------
# The fused step (afqmc_gto.py:630-725)
# ---------------------------------------------------------------------------

class System:
    """An ``afq_system``: integrals and trial prepared for one precision."""

    def __init__(self, propagator, chol, rchol_a, rchol_b, trial_up, trial_dn, fbbound=None,
                 exp_nmax=6, precision="fp64", chol_sign=None):
        self.chol = _r(chol)
        self.rchol_a = _r(rchol_a)
        self.rchol_b = _r(rchol_b)
        if self.chol.shape[0] == 0:
            # no Cholesky vectors (no two-body term): the ABI still takes
            # non-NULL arrays, which are then never read
            self._empty = torch.zeros(1, dtype=F64, device=self.chol.device)
        E = propagator['expH1']
        if isinstance(E, torch.Tensor) and E.is_complex():
            if float(torch.max(torch.abs(E.imag))) != 0.0:
                raise ValueError("expH1 must be real (mf_shift purely imaginary)")
            E = E.real
        self.expH1 = _r(E)
        self.mf_shift = _c(propagator['mf_shift'])
        self.trial_up = _c(trial_up)
        self.trial_dn = _c(trial_dn)
        self.dt = float(propagator['dt'])
        self.fbbound = FBBOUND_DEFAULT if fbbound is None else float(fbbound)
        self.nmax = int(exp_nmax)
        self.precision = precision
        G, M, _ = self.chol.shape
        self.M, self.G = M, G
        self.Na, self.Nb = self.trial_up.shape[1], self.trial_dn.shape[1]
        h = _lib._P()
        nz = (lambda t: ptr(self._empty) if t.numel() == 0 else ptr(t)) if G == 0 else ptr
        check(_L().afq_system_create(_ctx(), _lib.ctypes.byref(h), M, self.Na, self.Nb, G,
                                     nz(self.chol), nz(self.rchol_a), nz(self.rchol_b),
                                     ptr(self.expH1), nz(self.mf_shift), ptr(self.trial_up),
                                     ptr(self.trial_dn), self.dt, self.fbbound, self.nmax,
                                     _lib.PRECISIONS[precision]), "afq_system_create")
        self.handle = h
        self.chol_sign = None
        if chol_sign is not None and G > 0:
            # afqmc_pw_heg.py:161-254: ERI sign of each vector (HS phase i / 1)
            self.chol_sign = _r(chol_sign).reshape(G)
            torch.cuda.synchronize()
            check(_L().afq_system_set_chol_sign(h, ptr(self.chol_sign)),
                  "afq_system_set_chol_sign")

    def step(self, phia, phib, weights, e_hybrid, eshift, xi=None, key=None):
        """In-place step; returns overlap_new."""
        _ctx()
        W = phia.shape[0]
        ov = torch.empty((W,), dtype=C128, device=phia.device)
        seed = 0 if key is None else int(key.seed) & (2 ** 64 - 1)
        step = 0 if key is None else int(key.step)
        xi_t = _r(xi) if xi is not None else None
        if xi is None and key is None:
            raise ValueError("either xi or key is required")
        check(_L().afq_propagate_walkers(self.handle, W, ptr(phia), ptr(phib), ptr(weights),
                                         ptr(e_hybrid), float(eshift), ptr(xi_t), seed, step,
                                         ptr(ov)), "afq_propagate_walkers")
        return ov

    def set_propagator_block(self, cols):
        """Widest column block of the BF16x3 propagator (16, 32 or 48; other
        precisions ignore it): a tuning knob, the results do not change."""
        check(_L().afq_system_set_propagator_block(self.handle, int(cols)),
              "afq_system_set_propagator_block")

    def local_energy(self, phia, phib, h1e, enuc):
        _ctx()
        W = phia.shape[0]
        out = [torch.empty((W,), dtype=C128, device=phia.device) for _ in range(3)]
        h = _r(h1e)
        check(_L().afq_system_local_energy(self.handle, W, ptr(phia), ptr(phib), ptr(h),
                                           float(enuc), ptr(out[0]), ptr(out[1]), ptr(out[2])),
              "afq_system_local_energy")
        return tuple(out)

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and _lib._lib is not None:
                _lib._lib.afq_system_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


SYSTEM_CACHE_SIZE = 4


def _cached_system(propagator, arrays, params, make):
    """The prepared system of ``propagator`` for these input arrays.

    Entries are keyed by the identity of the caller's arrays and keep strong
    references to them, so an id cannot be recycled by another array while
    its entry lives; at most SYSTEM_CACHE_SIZE entries (least recently used
    first out) are kept per propagator, each holding device operands."""
    if not isinstance(propagator, dict):
        return make()
    key = tuple(id(a) for a in arrays) + tuple(params)
    cache = propagator.setdefault('_systems', {})
    hit = cache.pop(key, None)
    if hit is not None and all(x is y for x, y in zip(hit[0], arrays)):
        cache[key] = hit  # most recent last
        return hit[1]
    sysobj = make()
    cache[key] = (tuple(arrays), sysobj)
    while len(cache) > SYSTEM_CACHE_SIZE:
        cache.pop(next(iter(cache)))
    return sysobj


def _system_for(propagator, chol, rchol_a, rchol_b, trial_up, trial_dn, fbbound, exp_nmax,
                precision="fp64"):
    return _cached_system(
        propagator, (chol, rchol_a, rchol_b, trial_up, trial_dn),
        (fbbound, exp_nmax, precision),
        lambda: System(propagator, chol, rchol_a, rchol_b, trial_up, trial_dn, fbbound,
                       exp_nmax, precision))


def propagate_walkers(phia, phib, weights, overlap, e_hybrid, propagator, chol, rchol_a, rchol_b,
                      trial_up, trial_dn, eshift, rng_key, fbbound=None, exp_nmax=6,
                      walker_chunk_size=None, chol_chunk_g=None, xi=None, precision="fp64"):
    """afqmc_gto.py:630-725 -> (phia, phib, weights_new, overlap_new, e_hybrid_new, rng_key).

    ``overlap`` is ignored as in the reference.  ``rng_key`` is a PhiloxKey;
    its step's fields are used unless ``xi`` (W, G) is given.  Returns new
    arrays (inputs are not modified).
    """
    sysobj = _system_for(propagator, chol, rchol_a, rchol_b, trial_up, trial_dn, fbbound,
                         exp_nmax, precision)
    pa, pb = _c(phia).clone(), _c(phib).clone()
    w = _r(weights).clone()
    eh = _c(e_hybrid).clone()
    if rng_key is None:
        rng_key = PhiloxKey()
    next_key, key = rng_key.split() if isinstance(rng_key, PhiloxKey) else (rng_key, None)
    ov = sysobj.step(pa, pb, w, eh, eshift, xi=xi, key=key)
    return pa, pb, w, ov, eh, next_key


# ---------------------------------------------------------------------------
# Multi-determinant trials (greens_th4.py, energy_th4.py, afqmc_gto.py:728-839)
# ---------------------------------------------------------------------------

def occupations_from_trials(trials):
    """(ndet, M, N) dets built as eye(M)[:, occ] (afqmc_gto.py:1172-1174) ->
    occ (ndet, N) int32 on the device.  Raises ValueError for a det that is
    not a selection of unit columns (only occupation-string dets are built)."""
    t = trials.detach().cpu() if isinstance(trials, torch.Tensor) else torch.from_numpy(
        np.asarray(trials))
    t = t.to(torch.complex128) if not t.is_complex() else t
    nd, M, N = t.shape
    occ = torch.argmax(torch.abs(t), dim=1).to(torch.int32)  # (nd, N)
    ref = torch.zeros_like(t)
    if N:
        ref.scatter_(1, occ.long().unsqueeze(1), 1.0)
    if not torch.equal(t, ref):
        raise ValueError("multi-determinant trials must be columns of the identity "
                         "(occupation strings)")
    return occ.to(_dev()).contiguous()


def _occ(x):
    """Occupation strings given directly (int) or as dense trials."""
    if isinstance(x, (np.ndarray, torch.Tensor)) and not (
            (isinstance(x, np.ndarray) and np.issubdtype(x.dtype, np.integer))
            or (isinstance(x, torch.Tensor) and not x.is_floating_point() and not x.is_complex())):
        return occupations_from_trials(x)
    t = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    return t.to(device=_dev(), dtype=torch.int32).contiguous()


def greens_function_multidet(phia, phib, trials_up, trials_dn, ci_coeffs, det_chunk_size=5):
    """greens_th4.py:75-148 -> (Ga, Gb, overlap)."""
    pa, pb, ci = _c(phia), _c(phib), _c(ci_coeffs)
    oa, ob = _occ(trials_up), _occ(trials_dn)
    W, M, Na = pa.shape
    Nb, nd = pb.shape[2], ci.shape[0]
    Ga = torch.empty((W, M, M), dtype=C128, device=pa.device)
    Gb = torch.empty_like(Ga)
    ov = torch.empty((W,), dtype=C128, device=pa.device)
    check(_L().afq_greens_function_multidet(_ctx(), W, M, Na, Nb, nd, ptr(pa), ptr(pb), ptr(oa),
                                            ptr(ob), ptr(ci), int(det_chunk_size), ptr(Ga),
                                            ptr(Gb), ptr(ov)), "afq_greens_function_multidet")
    return Ga, Gb, ov


def greens_function_multidet_overlap_only(phia, phib, trials_up, trials_dn, ci_coeffs,
                                          det_chunk_size=5):
    """greens_th4.py:186-228 -> overlap."""
    pa, pb, ci = _c(phia), _c(phib), _c(ci_coeffs)
    oa, ob = _occ(trials_up), _occ(trials_dn)
    W, M, Na = pa.shape
    ov = torch.empty((W,), dtype=C128, device=pa.device)
    check(_L().afq_greens_function_multidet_overlap_only(
        _ctx(), W, M, Na, pb.shape[2], ci.shape[0], ptr(pa), ptr(pb), ptr(oa), ptr(ob), ptr(ci),
        int(det_chunk_size), ptr(ov)), "afq_greens_function_multidet_overlap_only")
    return ov


def local_energy_multidet(h1e, chol, phia, phib, trials_up, trials_dn, ci_coeffs, enuc,
                          det_chunk_size=5):
    """energy_th4.py:26-119 -> (e_tot, e_1b, e_2b)."""
    pa, pb, ci = _c(phia), _c(phib), _c(ci_coeffs)
    h, L = _r(h1e), _r(chol)
    oa, ob = _occ(trials_up), _occ(trials_dn)
    W, M, Na = pa.shape
    out = [torch.empty((W,), dtype=C128, device=pa.device) for _ in range(3)]
    check(_L().afq_local_energy_multidet(_ctx(), W, M, Na, pb.shape[2], L.shape[0], ci.shape[0],
                                         ptr(h), ptr(L), ptr(pa), ptr(pb), ptr(oa), ptr(ob),
                                         ptr(ci), float(enuc), int(det_chunk_size), ptr(out[0]),
                                         ptr(out[1]), ptr(out[2])), "afq_local_energy_multidet")
    return tuple(out)


def half_rotate_cholesky_multidet(chol, trials_up, trials_dn, chunk_g=None):
    """cholesky.py:513-560 -> rchols_a (ndet, G, Na, M), rchols_b (real for
    real trials; torch on the device, set-up only)."""
    L = _r(chol)
    out = []
    for T in (trials_up, trials_dn):
        T = T if isinstance(T, torch.Tensor) else torch.from_numpy(np.asarray(T))
        T = T.to(device=L.device)
        Tc = T.conj() if T.is_complex() else T.to(F64)
        if Tc.is_complex():
            r = torch.einsum('dpi,gpq->dgiq', Tc, L.to(C128))
        else:
            r = torch.einsum('dpi,gpq->dgiq', Tc, L)
        out.append(r.contiguous())
    return out[0], out[1]


def greens_function_multidet_force_bias(phia, phib, trials_up, trials_dn, ci_coeffs):
    """greens.py:314-369 -> (Ghalfa_all, Ghalfb_all, overlap, ovlp_a_all, ovlp_b_all)."""
    pa, pb, ci = _c(phia), _c(phib), _c(ci_coeffs)
    oa, ob = _occ(trials_up), _occ(trials_dn)
    W, M, Na = pa.shape
    Nb, nd = pb.shape[2], ci.shape[0]
    dev = pa.device
    Gha = torch.empty((nd, W, Na, M), dtype=C128, device=dev)
    Ghb = torch.empty((nd, W, Nb, M), dtype=C128, device=dev)
    ov = torch.empty((W,), dtype=C128, device=dev)
    ova = torch.empty((nd, W), dtype=C128, device=dev)
    ovb = torch.empty((nd, W), dtype=C128, device=dev)
    check(_L().afq_greens_function_multidet_force_bias(
        _ctx(), W, M, Na, Nb, nd, ptr(pa), ptr(pb), ptr(oa), ptr(ob), ptr(ci), ptr(Gha),
        ptr(Ghb), ptr(ov), ptr(ova), ptr(ovb)), "afq_greens_function_multidet_force_bias")
    return Gha, Ghb, ov, ova, ovb


def local_energy_multidet_streamed(h1e_trials_a, h1e_trials_b, Ghalfa_all, Ghalfb_all, rchols_a,
                                   rchols_b, ci_coeffs, ovlp_a_all, ovlp_b_all, enuc,
                                   e_chunk_g=16):
    """energy.py:318-395 -> (e_tot, e_1b, e_2b), no (W, M, M) Green's function."""
    hta, htb = _c(h1e_trials_a), _c(h1e_trials_b)
    Gha, Ghb = _c(Ghalfa_all), _c(Ghalfb_all)
    ra, dta = _real_or_complex(rchols_a)
    rb, dtb = _real_or_complex(rchols_b)
    if dta != dtb:
        ra, rb, dta = _c(ra), _c(rb), _lib.AFQ_COMPLEX
    ci, oa, ob = _c(ci_coeffs), _c(ovlp_a_all), _c(ovlp_b_all)
    nd, W, Na, M = Gha.shape
    Nb = Ghb.shape[2]
    G = ra.shape[1]
    out = [torch.empty((W,), dtype=C128, device=Gha.device) for _ in range(3)]
    check(_L().afq_local_energy_multidet_streamed(
        _ctx(), W, M, Na, Nb, G, nd, ptr(hta), ptr(htb), ptr(Gha), ptr(Ghb), ptr(ra), ptr(rb), dta,
        ptr(ci), ptr(oa), ptr(ob), float(enuc), int(e_chunk_g), ptr(out[0]), ptr(out[1]),
        ptr(out[2])), "afq_local_energy_multidet_streamed")
    return tuple(out)


def multidet_charge_density(trials_up, trials_dn, ci_coeffs):
    """afqmc_gto.py:1209-1218: sum_d |c_d|^2 (T_d T_d^H)_up + (...)_dn / sum |c|^2."""
    tu, td, ci = _c(trials_up), _c(trials_dn), _c(ci_coeffs)
    w = torch.abs(ci) ** 2
    w = (w / torch.sum(w)).to(C128)
    G = torch.einsum('d,dpi,dqi->pq', w, tu, tu.conj())
    return G + torch.einsum('d,dpi,dqi->pq', w, td, td.conj())


class MultiDetSystem(System):
    """An ``afq_system`` with a multi-determinant trial of occupation strings."""

    METHODS = {"explicit": 0, "ratio": 1, "wick": 1}

    def __init__(self, propagator, chol, trials_up, trials_dn, ci_coeffs, fbbound=None,
                 exp_nmax=6, det_chunk_size=5, precision="fp64", method=None):
        self.chol = _r(chol)
        E = propagator['expH1']
        if isinstance(E, torch.Tensor) and E.is_complex():
            if float(torch.max(torch.abs(E.imag))) != 0.0:
                raise ValueError("expH1 must be real (mf_shift purely imaginary)")
            E = E.real
        self.expH1 = _r(E)
        self.mf_shift = _c(propagator['mf_shift'])
        self.occ_up, self.occ_dn = _occ(trials_up), _occ(trials_dn)
        self.ci = _c(ci_coeffs)
        self.dt = float(propagator['dt'])
        self.fbbound = FBBOUND_DEFAULT if fbbound is None else float(fbbound)
        self.nmax = int(exp_nmax)
        self.precision = precision
        G, M, _ = self.chol.shape
        self.M, self.G = M, G
        self.Na, self.Nb = self.occ_up.shape[1], self.occ_dn.shape[1]
        self.ndet = self.ci.shape[0]
        eye = torch.eye(M, dtype=C128, device=self.chol.device)
        self.trial_up = eye[:, self.occ_up[0].long()].contiguous()   # dominant det
        self.trial_dn = eye[:, self.occ_dn[0].long()].contiguous()
        h = _lib._P()
        check(_L().afq_system_create_multidet(
            _ctx(), _lib.ctypes.byref(h), M, self.Na, self.Nb, G, ptr(self.chol),
            ptr(self.expH1), ptr(self.mf_shift), self.ndet, ptr(self.occ_up), ptr(self.occ_dn),
            ptr(self.ci), self.dt, self.fbbound, self.nmax, int(det_chunk_size),
            _lib.PRECISIONS[precision]), "afq_system_create_multidet")
        self.handle = h
        if method is not None:
            # "explicit": one inverse per (walker, det); "ratio" (or the
            # reference's name "wick"): ratio updates relative to det 0
            check(_L().afq_system_set_multidet_method(h, self.METHODS[method]),
                  "afq_system_set_multidet_method")

    @property
    def method(self):
        return {0: "explicit", 1: "ratio"}[int(_L().afq_system_get_multidet_method(self.handle))]


def propagate_walkers_multidet(phia, phib, weights, overlap, e_hybrid, propagator, chol,
                               trials_up, trials_dn, ci_coeffs, eshift, rng_key, fbbound=None,
                               exp_nmax=6, walker_chunk_size=None, chol_chunk_g=None,
                               det_chunk_size=5, xi=None, precision="fp64", method=None):
    """afqmc_gto.py:728-839 -> (phia, phib, weights_new, overlap_new, e_hybrid_new, rng_key).

    ``method``: None (ratio updates when every det is within four orbitals of
    det 0, else explicit), "explicit" or "ratio"."""
    sysobj = _cached_system(
        propagator, (chol, trials_up, trials_dn, ci_coeffs),
        (fbbound, exp_nmax, det_chunk_size, precision, "md", method),
        lambda: MultiDetSystem(propagator, chol, trials_up, trials_dn, ci_coeffs, fbbound,
                               exp_nmax, det_chunk_size, precision, method))
    pa, pb = _c(phia).clone(), _c(phib).clone()
    w = _r(weights).clone()
    eh = _c(e_hybrid).clone()
    if rng_key is None:
        rng_key = PhiloxKey()
    next_key, k = rng_key.split() if isinstance(rng_key, PhiloxKey) else (rng_key, None)
    ov = sysobj.step(pa, pb, w, eh, eshift, xi=xi, key=k)
    return pa, pb, w, ov, eh, next_key


class WickTrial:
    """The multi-det trial of ``build_wick_trial`` (wick.py:190-226), kept as
    occupation strings and CI coefficients: the ratio-update step forms what
    the reference's WickTrial precomputes (det 0's inverse, each det's
    excitation, the adjugate weights) on the device at every step."""

    def __init__(self, occ_up, occ_dn, ci_coeffs):
        self.occ_up = occ_up
        self.occ_dn = occ_dn
        self.ci_coeffs = ci_coeffs
        self.ndet = len(ci_coeffs)


def build_wick_trial(occ_up, occ_dn, ci_coeffs, chol=None, h1e=None, chunk_g=None, orbs_up=None,
                     orbs_dn=None):
    """wick.py:190-226 for occupation-string trials in the walker basis (det 0
    the reference determinant).  Other orbital sets (UHF-based trials) are
    not built."""
    if orbs_up is not None or orbs_dn is not None:
        raise NotImplementedError("Wick trials over other orbital sets are not built")
    return WickTrial(occ_up, occ_dn, ci_coeffs)


def propagate_walkers_wick(phia, phib, weights, overlap, e_hybrid, propagator, chol, wick_trial,
                           eshift, rng_key, fbbound=None, exp_nmax=6, walker_chunk_size=None,
                           chol_chunk_g=None, det_chunk_size=5, xi=None, precision="fp64"):
    """afqmc_gto.py:842-935 -> (phia, phib, weights_new, overlap_new, e_hybrid_new, rng_key):
    the multi-det step on ratio updates relative to det 0 (one inverse per
    walker and spin)."""
    t = wick_trial
    return propagate_walkers_multidet(phia, phib, weights, overlap, e_hybrid, propagator, chol,
                                      t.occ_up, t.occ_dn, t.ci_coeffs, eshift, rng_key, fbbound,
                                      exp_nmax, walker_chunk_size, chol_chunk_g, det_chunk_size,
                                      xi=xi, precision=precision, method="ratio")


# ---------------------------------------------------------------------------
# Per-walker numerical status (no reference counterpart: the reference lets
# NaN/inf propagate and masks only multi-det overlaps, greens.py:264-276)
# ---------------------------------------------------------------------------

def walker_status(phia, phib, overlap=None, weights=None, e_hybrid=None, zero_overlap_tol=1e-300):
    """Per-walker flag words (int32 tensor (W,), bits of ``_lib.WALKER_FLAGS``)
    and a device summary int32 (2,): [OR of all flags, number of walkers with
    a flag other than zero_weight]."""
    pa, pb = _c(phia), _c(phib)
    W, M, Na = pa.shape
    Nb = pb.shape[2]
    ov = _c(overlap) if overlap is not None else None
    w = _r(weights) if weights is not None else None
    eh = _c(e_hybrid) if e_hybrid is not None else None
    flags = torch.empty((W,), dtype=torch.int32, device=pa.device)
    summary = torch.empty((2,), dtype=torch.int32, device=pa.device)
    check(_L().afq_walker_status(_ctx(), W, M, Na, Nb, ptr(pa), ptr(pb), ptr(ov), ptr(w), ptr(eh),
                                 float(zero_overlap_tol), ptr(flags), ptr(summary)),
          "walker_status")
    return flags, summary


def decode_walker_flags(word):
    """Names of the bits set in one flag word."""
    return [k for k, v in _lib.WALKER_FLAGS.items() if int(word) & v]


# ---------------------------------------------------------------------------
# Statistics (utils.py:86-149)
# ---------------------------------------------------------------------------

def do_binning_analysis(a):
    """utils.py:86-149 -> (mean, serr, std, kappa) of a short host series."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    xbar = a.mean()
    sdev = a.std()
    x = a - xbar
    ac = np.correlate(x, x, mode='full')
    with np.errstate(invalid='ignore', divide='ignore'):
        # a constant series gives 0/0: no positive lag then, kappa = 1
        ac = ac / ac[ac.shape[0] // 2]
    kappa = 1.0
    i = (ac.shape[0] >> 1) + 1
    while i < ac.shape[0] and ac[i] >= 0:
        kappa += 2.0 * ac[i]
        i += 1
    return xbar, sdev * (kappa / n) ** 0.5, sdev, kappa
