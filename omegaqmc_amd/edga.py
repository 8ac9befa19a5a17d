"""EDGA multi-determinant trial files -> the multi-det trial dict.

Mirrors ``examples/dmrg_afqmc/edga2omegaqmc.py`` (``load_edga_result``,
``make_omegaqmc_trial``) and the determinant selection of
``examples/dmrg_afqmc/test_h2_afqmc.py:248-253``, so that a trial written by
EDGA reaches :class:`omegaqmc_amd.driver.AFQMCDriver` (``trial=...``) the way
it reaches the reference's ``get_afqmc_func``.

File format (text)::

    # EDGA: state=0 mode=genetic completeness=0.99 n_det=11857 gen=27
      -0.518753925297355  4444411111  |0 1 2 3 4  |0 1 2 3 4
       0.246489595130581  1444441111  |1 2 3 4 5  |1 2 3 4 5

one determinant per line: CI coefficient, occupation string over the active
orbitals (1 empty, 2 alpha, 3 beta, 4 both), then the alpha and the beta
orbital indices, each list after a ``|``.  Header lines start with ``#``;
their ``key=value`` tokens become the metadata.
"""

import numpy as np

__all__ = ["load_edga_result", "make_omegaqmc_trial", "select_top_determinants",
           "occupation_lists"]


def _header_tokens(line, meta):
    for tok in line.split():
        key, eq, val = tok.partition('=')
        if eq:
            meta[key] = val


def load_edga_result(filename, coeff_threshold=0.0):
    """edga2omegaqmc.py:46-101 -> (dets, coeffs, meta).

    ``dets[k]`` is ``{'occ': str, 'alpha': [int], 'beta': [int]}``; a
    determinant with ``|c| < coeff_threshold`` is dropped; lines with fewer
    than two ``|`` separators are skipped, as in the reference.
    """
    dets, coeffs, meta = [], [], {}
    with open(filename, 'r') as fh:
        for raw in fh:
            line = raw.strip()
            if not line:
                continue
            if line.startswith('#'):
                _header_tokens(line, meta)
                continue
            fields = line.split('|')
            if len(fields) < 3:
                continue
            head = fields[0].split()
            c = float(head[0])
            if abs(c) < coeff_threshold:
                continue
            coeffs.append(c)
            dets.append({'occ': head[1],
                         'alpha': [int(x) for x in fields[1].split()],
                         'beta': [int(x) for x in fields[2].split()]})
    return dets, coeffs, meta


def occupation_lists(occ):
    """Alpha and beta orbital indices of an occupation string (1 empty,
    2 alpha, 3 beta, 4 both)."""
    alpha = [k for k, ch in enumerate(occ) if ch in '24']
    beta = [k for k, ch in enumerate(occ) if ch in '34']
    return alpha, beta


def select_top_determinants(dets, coeffs, max_det):
    """The ``max_det`` determinants of largest |c| in the order
    ``np.argsort(-|c|)`` gives (test_h2_afqmc.py:250-253); all of them, in
    file order, when ``max_det`` is falsy or not smaller than their number."""
    if not max_det or len(dets) <= max_det:
        return list(dets), list(coeffs)
    order = np.argsort(-np.abs(np.asarray(coeffs, dtype=np.float64)))[:max_det]
    return [dets[k] for k in order], [coeffs[k] for k in order]


def make_omegaqmc_trial(coeffs, dets, mo_coeff, n_core=0, index_map=None, meta=None):
    """edga2omegaqmc.py:104-175 -> trial dict.

    Keys ``ci_coeffs`` (float64, (ndet,)), ``occ_up`` / ``occ_dn`` (int32,
    (ndet, n_core + n_active)), ``ndet`` and ``mo_coeff``.  EDGA indices are
    mapped by ``index_map`` when given (ignored when ``meta['indexing']`` is
    ``'active'``), else shifted by ``n_core``; the core orbitals
    0 .. n_core-1 are added to every determinant and each row is sorted.
    """
    if meta and meta.get('indexing') == 'active':
        if index_map is not None:
            print("EDGA header says indexing=active: the given index_map is not used")
        index_map = None
    ndet = len(coeffs)
    na, nb = len(dets[0]['alpha']), len(dets[0]['beta'])
    core = list(range(n_core))
    occ_up = np.zeros((ndet, n_core + na), dtype=np.int32)
    occ_dn = np.zeros((ndet, n_core + nb), dtype=np.int32)

    def mapped(idx):
        if index_map is not None:
            return [index_map[k] for k in idx]
        return [k + n_core for k in idx] if n_core > 0 else list(idx)

    for k, det in enumerate(dets):
        a, b = det['alpha'], det['beta']
        if len(a) != na or len(b) != nb:
            raise ValueError(f"determinant {k} has {len(a)} up / {len(b)} down orbitals, "
                             f"the first one {na} / {nb}")
        occ_up[k] = sorted(core + mapped(a))
        occ_dn[k] = sorted(core + mapped(b))
    return {'ci_coeffs': np.asarray(coeffs, dtype=np.float64), 'occ_up': occ_up,
            'occ_dn': occ_dn, 'ndet': ndet, 'mo_coeff': np.asarray(mo_coeff)}
