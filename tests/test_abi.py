"""CPU checks of the C ABI: the library builds, loads without a GPU, exports
every function include/qtip4pf.h declares, and host-only pieces (Philox block
function, reference-order host draws) behave."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "qtip4pf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(qtp_[a-z0-9_]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from qtip4pf_water_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from qtip4pf_water_amd import build
        build.build()
    return _lib.load()


def test_header_parses():
    names = header_functions()
    assert "qtp_create" in names and "qtp_update_sc_nospline" in names
    assert len(names) >= 25


def test_every_declared_symbol_is_exported(lib):
    from qtip4pf_water_amd import _lib
    bound = {n for n, _, _ in _lib.SIGNATURES}
    for name in header_functions():
        assert hasattr(lib, name), name
        assert name in bound, f"{name} has no ctypes signature"


def test_philox_known_answers(lib):
    # Random123 known-answer vectors for philox4x32_10
    from qtip4pf_water_amd import philox4x32_10
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = philox4x32_10(ctr, key)
        assert tuple(int(x) for x in got) == want


def test_bad_params_rejected_without_gpu(lib):
    from qtip4pf_water_amd import Water, QtpError
    with pytest.raises(QtpError):
        Water(216, 3, (35.0, 35.0, 35.0))   # odd bead count (water.cpp:69-72)


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="oracle/_ref not built")
def test_host_draws_follow_reference_order(lib, oracle_mod):
    """The library's std::mt19937 draws equal the reference harness's draws."""
    from qtip4pf_water_amd import HostRng
    o = oracle_mod.Oracle(8, 4, (12.0, 12.0, 12.0))
    oracle_mod.ref_set_from(o)
    R = oracle_mod.ref()
    for seed in (1, 77):
        R.ref_seed(seed)
        want = np.zeros(o.n_deviates())
        R.ref_draw_pile_deviates(oracle_mod._ptr(want))
        got = HostRng(seed).draw_pile(o.n_atoms, o.n_beads)
        assert np.array_equal(got, want)
