"""EDGA multi-determinant trial reader (omegaqmc_amd/edga.py) on the
reference's own example trial (tests/golden/edga_result_head.dat: the first
300 of the 11857 determinants of examples/dmrg_afqmc/edga_result.dat)."""

import pathlib

import numpy as np

from omegaqmc_amd import edga as E

HEAD = pathlib.Path(__file__).with_name('golden') / 'edga_result_head.dat'


def test_parse_reference_head():
    dets, coeffs, meta = E.load_edga_result(HEAD)
    assert meta == {'state': '0', 'mode': 'genetic', 'completeness': '0.9900075036',
                    'n_det': '11857', 'gen': '27'}
    assert len(dets) == len(coeffs) == 300
    assert coeffs[0] == -0.518753925297355 and dets[0]['occ'] == '4444411111'
    assert dets[0]['alpha'] == dets[0]['beta'] == [0, 1, 2, 3, 4]
    # the orbital lists restate the occupation string (1 empty, 2 alpha,
    # 3 beta, 4 both): an independent check of the column split
    for d in dets:
        assert (d['alpha'], d['beta']) == E.occupation_lists(d['occ'])
        assert len(d['occ']) == 10 and len(d['alpha']) == len(d['beta']) == 5
    # the file lists determinants by decreasing |c|
    a = np.abs(np.array(coeffs))
    assert np.all(np.diff(a) <= 0)


def test_threshold_and_top_selection():
    dets, coeffs, _ = E.load_edga_result(HEAD)
    kd, kc, _ = E.load_edga_result(HEAD, coeff_threshold=0.02)
    keep = [k for k, c in enumerate(coeffs) if abs(c) >= 0.02]
    assert kc == [coeffs[k] for k in keep] and kd == [dets[k] for k in keep]
    assert 0 < len(kc) < len(coeffs)
    sd, sc = E.select_top_determinants(dets[::-1], coeffs[::-1], 40)
    assert sorted(np.abs(sc).tolist(), reverse=True) == np.abs(coeffs[:40]).tolist()
    assert all(d == dets[coeffs.index(c)] for d, c in zip(sd, sc))
    assert E.select_top_determinants(dets, coeffs, None) == (dets, coeffs)
    assert E.select_top_determinants(dets, coeffs, 1000) == (dets, coeffs)


def test_trial_core_shift_and_maps():
    dets, coeffs, _ = E.load_edga_result(HEAD)
    mo = np.eye(13)
    t = E.make_omegaqmc_trial(coeffs[:20], dets[:20], mo, n_core=3)
    assert t['ndet'] == 20 and t['occ_up'].shape == (20, 8) and t['occ_dn'].shape == (20, 8)
    assert t['occ_up'].dtype == np.int32 and t['ci_coeffs'].dtype == np.float64
    for k in range(20):
        assert t['occ_up'][k].tolist() == [0, 1, 2] + [a + 3 for a in dets[k]['alpha']]
        assert t['occ_dn'][k].tolist() == [0, 1, 2] + [b + 3 for b in dets[k]['beta']]
    np.testing.assert_array_equal(t['ci_coeffs'], coeffs[:20])
    perm = [4, 8, 3, 2, 7, 1, 6, 0, 5, 9]
    tm = E.make_omegaqmc_trial(coeffs[:20], dets[:20], mo, index_map=perm)
    for k in range(20):
        assert tm['occ_up'][k].tolist() == sorted(perm[a] for a in dets[k]['alpha'])
    # indexing=active: the map is ignored, only the core shift applies
    ta = E.make_omegaqmc_trial(coeffs[:20], dets[:20], mo, n_core=3, index_map=perm,
                               meta={'indexing': 'active'})
    np.testing.assert_array_equal(ta['occ_up'], t['occ_up'])


def test_malformed_lines_skipped(tmp_path):
    f = tmp_path / 'e.dat'
    f.write_text("# EDGA: state=1 n_det=2\n\n 0.5 4211 |0 1 |0\n bad line\n"
                 "-0.25 2411 |0 1|1\n  0.1 4411 |0 1\n")
    dets, coeffs, meta = E.load_edga_result(f)
    assert meta == {'state': '1', 'n_det': '2'}
    assert coeffs == [0.5, -0.25]
    assert [(d['alpha'], d['beta']) for d in dets] == [([0, 1], [0]), ([0, 1], [1])]
