"""Brute-force second quantisation for tiny systems (test helper, host only).

Builds one-body operators E_pq = a_p^+ a_q on the N-electron determinant
space of each spin (Jordan-Wigner signs on bit strings), the full
Hamiltonian H = sum h'_pq E_pq + 1/2 sum_g (sum_pq L_gpq E_pq)^2 + E_nuc
with h' = h - 1/2 sum_g L_g L_g (the form AFQMC factorises), and CI
vectors of Slater determinants (coefficients are N x N minors).  Used to
pin the oracle's overlap and mixed local energy independently of its own
formulas: E_mixed(phi) = <Psi_T|H|phi> / <Psi_T|phi>.
"""

from __future__ import annotations

import itertools

import numpy as np


def strings(M, N):
    return [sum(1 << i for i in c) for c in itertools.combinations(range(M), N)]


def _popcount_below(s, i):
    return bin(s & ((1 << i) - 1)).count("1")


def one_body_ops(M, N):
    S = strings(M, N)
    idx = {s: i for i, s in enumerate(S)}
    D = len(S)
    E = np.zeros((M, M, D, D))
    for j, s in enumerate(S):
        for q in range(M):
            if not (s >> q) & 1:
                continue
            s1 = s ^ (1 << q)
            sg1 = (-1) ** _popcount_below(s, q)
            for p in range(M):
                if (s1 >> p) & 1:
                    continue
                s2 = s1 | (1 << p)
                sg2 = (-1) ** _popcount_below(s1, p)
                E[p, q, idx[s2], j] += sg1 * sg2
    return E


def hamiltonian(h1e_mod, chol, enuc, M, na, nb):
    Ea, Eb = one_body_ops(M, na), one_body_ops(M, nb)
    Da, Db = Ea.shape[2], Eb.shape[2]

    def full(h):
        return (np.kron(np.einsum('pq,pqij->ij', h, Ea), np.eye(Db))
                + np.kron(np.eye(Da), np.einsum('pq,pqij->ij', h, Eb)))

    H = full(h1e_mod) + enuc * np.eye(Da * Db)
    for g in range(chol.shape[0]):
        V = full(chol[g])
        H = H + 0.5 * V @ V
    return H


def det_vector(phi):
    """CI coefficients of the determinant with orbitals phi (M x N) on strings(M, N)."""
    M, N = phi.shape
    out = []
    for c in itertools.combinations(range(M), N):
        out.append(np.linalg.det(phi[list(c), :]))
    return np.array(out)


def mixed_energy(H, trial_a, trial_b, phia, phib):
    """<T|H|phi>/<T|phi> and <T|phi> for one walker (phi_a x phi_b product)."""
    t = np.kron(det_vector(trial_a), det_vector(trial_b))
    v = np.kron(det_vector(phia), det_vector(phib))
    ov = np.vdot(t, v)
    return np.vdot(t, H @ v) / ov, ov
