"""The C-ABI library loads without a GPU and exports every declared symbol."""

import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        names |= set(re.findall(r"^\s*(?:int|double|const char\*)\s+(afq_\w+)\s*\(", txt, re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for must in ("afq_propagate_walkers", "afq_greens_function_force_bias", "afq_local_energy",
                 "afq_population_control_comb", "afq_orthogonalize_walkers", "afq_build_vhs"):
        assert must in names


def test_library_exports_every_declared_symbol(afq_lib):
    missing = [n for n in sorted(_declared()) if not hasattr(afq_lib, n)]
    assert not missing, missing


def test_binding_signatures_match_library(afq_lib):
    from omegaqmc_amd import _lib
    declared = _declared()
    assert set(_lib.SIGNATURES) <= declared
    for n in _lib.SIGNATURES:
        assert hasattr(afq_lib, n)


def test_host_only_entry_points_run_without_gpu(afq_lib):
    from oracle import philox as P
    for seed, step in ((42, 1), (7, 3), (2 ** 63 + 5, 99)):
        assert afq_lib.afq_comb_offset(seed, step) == P.comb_offset(seed, step)
    assert afq_lib.afq_version() >= 10000
    # invalid arguments are rejected before any device work
    rc = afq_lib.afq_build_vhs(None, 1, 1, 1, None, None, 0.005, None)
    assert rc == -1
    assert b"ctx" in afq_lib.afq_last_error()
