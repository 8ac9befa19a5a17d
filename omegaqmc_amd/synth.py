"""Synthetic AFQMC Hamiltonians (one-body h1, symmetric Cholesky vectors L_g).

The workloads this package is measured on are defined by a generator rather
than by PySCF molecules:

* ``h1`` is symmetric; its diagonal is ``linspace(-2, +1, M)`` Ha and its
  strict upper triangle is drawn N(0, 0.05) (standard deviation) in the
  row-major order of ``numpy.triu_indices(M, 1)`` and mirrored.
* Each Cholesky vector ``L_g`` is a symmetric M x M matrix whose upper
  triangle (diagonal included, row-major ``triu_indices(M)`` order) is drawn
  N(0, sigma_g) with standard deviation ``sigma_g = 0.3 * exp(-g / lam)`` and
  mirrored; the decay length ``lam`` defaults to ``naux / 2``.
* ``E_nuc = 1.0``.
* The trial is the RHF-like determinant built from the lowest ``N``
  eigenvectors ``C`` of ``h1``.  As in the reference array path
  (``OmegaQMC/integrals/cholesky.py:392-445``) every integral is rotated into
  that eigenbasis (``h1 -> C^T h1 C``, ``L_g -> C^T L_g C``) so the trial is
  ``eye(M, N)`` and ``h1e_mod = h1e + v0`` with
  ``v0 = -1/2 sum_gj L_gij L_gkj`` (``cholesky.py:261-275``).

Eigenvector signs are fixed (largest-magnitude component positive) so the
rotated integrals do not depend on the LAPACK build.

All draws come from ``numpy.random.default_rng(seed)`` in a fixed order
(h1 off-diagonals, then all L_g at once), so a (M, naux, seed) triple always
gives the same arrays.
"""

from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class SyntheticSystem:
    """Integrals in the trial (h1-eigen) basis, plus the raw draws."""

    nbasis: int
    nup: int
    ndown: int
    naux: int
    h1e: np.ndarray        # (M, M) rotated one-body Hamiltonian
    h1e_mod: np.ndarray    # (M, M) h1e + v0
    chol: np.ndarray       # (G, M, M) rotated, symmetric Cholesky vectors
    enuc: float
    mo_coeff: np.ndarray   # (M, M) rotation used (eigenvectors of raw h1)
    h1_raw: np.ndarray     # (M, M) unrotated h1
    decay: float

    @property
    def trial_up(self) -> np.ndarray:
        return np.eye(self.nbasis, self.nup)

    @property
    def trial_dn(self) -> np.ndarray:
        return np.eye(self.nbasis, self.ndown)


def _fix_signs(C: np.ndarray) -> np.ndarray:
    idx = np.argmax(np.abs(C), axis=0)
    s = np.sign(C[idx, np.arange(C.shape[1])])
    s[s == 0] = 1.0
    return C * s[None, :]


def raw_integrals(nbasis: int, naux: int, seed: int = 42, decay: float | None = None,
                  off_sigma: float = 0.05, chol_sigma: float = 0.3):
    """Draw the unrotated (h1, L) pair of the generator described above."""
    M, G = int(nbasis), int(naux)
    lam = float(decay) if decay is not None else G / 2.0
    rng = np.random.default_rng(seed)
    h1 = np.zeros((M, M))
    iu1 = np.triu_indices(M, 1)
    h1[iu1] = rng.normal(0.0, off_sigma, size=len(iu1[0]))
    h1 = h1 + h1.T
    h1[np.diag_indices(M)] = np.linspace(-2.0, 1.0, M)

    iu = np.triu_indices(M)
    sig = chol_sigma * np.exp(-np.arange(G) / lam)
    vals = rng.standard_normal((G, len(iu[0]))) * sig[:, None]
    L = np.zeros((G, M, M))
    L[:, iu[0], iu[1]] = vals
    L[:, iu[1], iu[0]] = vals
    return h1, L, lam


def one_body_chol_correction(chol: np.ndarray) -> np.ndarray:
    """v0[i,k] = -1/2 sum_{g,j} L[g,i,j] L[g,k,j] (cholesky.py:261-275)."""
    G, M = chol.shape[0], chol.shape[1]
    A = chol.transpose(0, 2, 1).reshape(G * M, M)
    return -0.5 * (A.T @ A)


REF_GM = 11 * 50  # G * M of the LiH-like workload


def size_normalised_sigma(nbasis: int, naux: int, chol_sigma: float = 0.3) -> float:
    """sigma_0 scaled by sqrt(550 / (G M)).

    With the literal sigma_0 = 0.3, ||VHS|| grows like sqrt(G M) and the
    one-body generator like G M: at M=300, G=1800 the determinant energy is
    -6.0e4 Ha, ||expH1|| = 2.4e7 and the walker overlap sign*exp(logdet)
    overflows FP64 in the first step of the reference algorithm.  Scaling the
    variance by 550/(G M) keeps ||VHS|| and the per-orbital two-body energy at
    their values for the LiH-like workload (where the factor is exactly 1).
    """
    return chol_sigma * float(np.sqrt(REF_GM / (int(nbasis) * int(naux))))


def make_system(nbasis: int, nup: int, ndown: int, naux: int, seed: int = 42,
                decay: float | None = None, enuc: float = 1.0,
                rotate_device: str | None = None,
                chol_sigma: float = 0.3, normalise: bool = False) -> SyntheticSystem:
    """Build a synthetic system, rotated into the basis of the trial.

    ``rotate_device`` (e.g. ``"cuda"``) performs the O(G M^3) basis rotation
    and the v0 contraction with torch on that device; the default runs them
    with NumPy/BLAS on the host.  ``normalise`` applies
    :func:`size_normalised_sigma` to ``chol_sigma``.
    """
    if normalise:
        chol_sigma = size_normalised_sigma(nbasis, naux, chol_sigma)
    h1, L, lam = raw_integrals(nbasis, naux, seed, decay, chol_sigma=chol_sigma)
    _, C = np.linalg.eigh(h1)
    C = _fix_signs(C)
    h1r = C.T @ h1 @ C
    h1r = 0.5 * (h1r + h1r.T)
    # The rotated L_g are symmetrised exactly ((L + L^T)/2 is bitwise
    # symmetric), which a rounded C^T L C is not.
    if rotate_device is None:
        Lr = np.matmul(np.matmul(C.T[None], L), C[None])
        Lr = 0.5 * (Lr + Lr.transpose(0, 2, 1))
        v0 = one_body_chol_correction(Lr)
    else:
        import torch
        dev = torch.device(rotate_device)
        Ct = torch.from_numpy(C).to(dev)
        Lt = torch.from_numpy(L).to(dev)
        Lt = torch.matmul(torch.matmul(Ct.T.unsqueeze(0), Lt), Ct.unsqueeze(0))
        Lt = 0.5 * (Lt + Lt.transpose(1, 2))
        A = Lt.transpose(1, 2).reshape(-1, nbasis)
        v0 = (-0.5 * (A.T @ A)).cpu().numpy()
        Lr = Lt.cpu().numpy()
        del Lt, A
    return SyntheticSystem(
        nbasis=int(nbasis), nup=int(nup), ndown=int(ndown), naux=int(naux),
        h1e=h1r, h1e_mod=h1r + v0, chol=np.ascontiguousarray(Lr), enuc=float(enuc),
        mo_coeff=C, h1_raw=h1, decay=lam)


# Workloads (M, nup, ndown, naux, walkers, steps).
CONFIGS = {
    "lih_like": dict(nbasis=11, nup=2, ndown=2, naux=50, walkers=256, steps=400),
    "m120": dict(nbasis=120, nup=16, ndown=16, naux=600, walkers=2048, steps=200),
    "m300": dict(nbasis=300, nup=40, ndown=40, naux=1800, walkers=1024, steps=100),
    "m150_open": dict(nbasis=150, nup=20, ndown=18, naux=800, walkers=1024, steps=100),
}


def multidet_trial(nbasis: int, nup: int, ndown: int, ndet: int = 32, seed: int = 11):
    """Multi-determinant trial of the M=150 workload: the aufbau determinant,
    then alpha singles, beta singles and alpha-beta doubles near the Fermi
    level (a deterministic list in the h1 eigenbasis), coefficients
    c_k = N(0, 1) exp(-k/4) with c_0 = 1, normalised.  Returns
    {'ndet', 'ci_coeffs', 'occ_up', 'occ_dn'} as the reference driver takes."""
    aufa, aufb = list(range(nup)), list(range(ndown))
    dets = [(tuple(aufa), tuple(aufb))]

    def single(occ, n, i, a):
        return tuple(sorted(set(occ) - {i} | {a}))

    k = 3
    sa = [(single(aufa, nup, i, a), tuple(aufb)) for i in range(nup - 1, max(-1, nup - 1 - k), -1)
          for a in range(nup, min(nbasis, nup + k))]
    sb = [(tuple(aufa), single(aufb, ndown, j, b)) for j in range(ndown - 1, max(-1, ndown - 1 - k), -1)
          for b in range(ndown, min(nbasis, ndown + k))]
    dab = [(x[0], y[1]) for x in sa for y in sb]
    for d in sa + sb + dab:
        if len(dets) >= ndet:
            break
        if d not in dets:
            dets.append(d)
    rng = np.random.default_rng(seed)
    ci = rng.normal(size=len(dets)) * np.exp(-np.arange(len(dets)) / 4.0)
    ci[0] = 1.0
    ci = ci / np.linalg.norm(ci)
    return {'ndet': len(dets), 'ci_coeffs': ci.astype(np.complex128),
            'occ_up': np.array([d[0] for d in dets], dtype=np.int32),
            'occ_dn': np.array([d[1] for d in dets], dtype=np.int32)}


def uneven_batch_specs(nsys: int = 64, seed: int = 7):
    """Systems of the uneven batch: M ~ U{40..200}, N = M // 8, G = 5 M,
    walkers from {128, 256, 512, 1024}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsys):
        M = int(rng.integers(40, 201))
        W = int(rng.choice([128, 256, 512, 1024]))
        out.append(dict(nbasis=M, nup=M // 8, ndown=M // 8, naux=5 * M, walkers=W))
    return out
