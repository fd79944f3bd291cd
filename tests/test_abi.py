"""CPU checks of the C ABI: the library builds, loads without a GPU, exports
every function include/qtip4pf.h declares, and host-only pieces (Philox block
function, reference-order host draws) behave."""
import math
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


def test_short_arrays_rejected_before_the_call(lib):
    """The C calls take bare pointers: every wrapper checks the length of each
    array it hands over (ValueError), so a short array never reaches them."""
    from qtip4pf_water_amd import Water
    from qtip4pf_water_amd import _lib
    w = Water.__new__(Water)          # sizes only; no device context is made
    w._L, w._h = _lib.load(), None
    w.n_mols, w.n_beads, w.list_size = 8, 4, 100
    w.n_atoms, w.n_dof = 24, 72
    P, nd = 4, 72
    q, F = np.zeros(P * nd), np.zeros(P * nd)
    chg, mass = np.zeros(24), np.zeros(72)
    point, lst = np.zeros(P * 8, np.int32), np.zeros(P * 100, np.int32)
    disp, dev = np.zeros(P * 24), np.zeros(3 + (P - 1) * nd)
    short = {"q": q[:-1], "F": F[:-1], "charge": chg[:-1], "point": point[:-1],
             "list": lst[:-1]}
    for name, bad in short.items():
        args = dict(q=q, F=F, charge=chg, point=point, lst=lst)
        args[{"list": "lst"}.get(name, name)] = bad
        with pytest.raises(ValueError, match=name):
            w.get_forces_nospline(args["q"], args["F"], args["charge"], args["point"], args["lst"])
    with pytest.raises(ValueError, match="fphys"):
        w.force_eval_sc_nospline(q, F, chg, mass, point, lst, fphys=np.zeros(P * nd - 3))
    with pytest.raises(ValueError, match="mass"):
        w.force_eval_sc_nospline(q, F, chg, mass[:-1], point, lst)
    with pytest.raises(ValueError, match="disp"):
        w.neighbour_list(q, point, lst, disp[:-1])
    with pytest.raises(ValueError, match="dev"):
        w.pile_g(q.copy(), mass, dev[:-1])
    with pytest.raises(ValueError, match="dev"):
        w.update_sc_nospline(q, q.copy(), mass, F, F.copy(), chg, point, lst, disp, dev[:-1],
                             True, 0.0)
    with pytest.raises(ValueError, match="p"):
        w.set_state(q, q[:-9])
    with pytest.raises(ValueError, match="a"):
        w.ftransform(np.zeros(P * 10 - 1), 10, True)
    with pytest.raises(ValueError, match="fphys"):
        w.get_kcv(q, F[:-1])


def test_kmax_ceiling_stated_and_enforced(lib):
    """k_max above what the LDS-resident factor tables allow is rejected at
    create (before any device work), with the ceiling in the message."""
    from qtip4pf_water_amd import QtpError, Water
    L = 100.0
    ew = dict(kappa=0.15, rcut=40.0, k_cut=3.0, k_max=46)   # kappa r_c = 6 (allowed)
    with pytest.raises(QtpError, match=r"k_max <= 45.*k_max = 46"):
        Water(64, 1, (L, L, L), ewald=ew)


def _philox_normal(seed, step, stream, idx):
    """Host mirror of the device thermostat's normal deviate: Philox4x32-10
    (the library's own block function) on counter (idx, step lo, stream,
    step hi) and key (seed lo, seed hi), two 53-bit uniforms, Box-Muller."""
    from qtip4pf_water_amd import philox4x32_10
    c = philox4x32_10((idx, step & 0xFFFFFFFF, stream, step >> 32),
                      (seed & 0xFFFFFFFF, seed >> 32)).astype(np.uint64)
    inv53 = 1.0 / 9007199254740992.0
    u1 = (float(((int(c[0]) >> 5) << 26) | (int(c[1]) >> 6)) + 0.5) * inv53
    u2 = (float(((int(c[2]) >> 5) << 26) | (int(c[3]) >> 6)) + 0.5) * inv53
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def test_thermostat_normals_moments_and_independence(lib):
    """The device thermostat's normals (mirrored on the host): mean 0,
    variance 1, kurtosis 3 over 40,000 draws, and no correlation between the
    streams of neighbouring modes, indices or steps (|r| < 0.02)."""
    seed = 12345
    n = 10000
    z = np.array([[_philox_normal(seed, s, k, i) for k in (1, 2)] for s in (7, 8)
                  for i in range(n // 2)])
    x = z.reshape(-1, 2)
    flat = x.ravel()
    assert abs(flat.mean()) < 0.03
    assert abs(flat.var() - 1.0) < 0.03
    kurt = np.mean(flat ** 4) / flat.var() ** 2
    assert abs(kurt - 3.0) < 0.15
    assert abs(np.corrcoef(x[:, 0], x[:, 1])[0, 1]) < 0.02      # stream k vs k + 1
    a = np.array([_philox_normal(seed, 7, 1, i) for i in range(n)])
    assert abs(np.corrcoef(a[:-1], a[1:])[0, 1]) < 0.02          # index i vs i + 1
    b = np.array([_philox_normal(seed, 8, 1, i) for i in range(n)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02                   # step vs step + 1


def test_destroy_frees_every_context_buffer():
    """Every device buffer the context allocates (dalloc(ctx, &ctx->X, ...)
    anywhere in the library's sources) is in qtp_destroy's free list, so a
    destroyed context returns all of its memory.  A source check: it needs
    no GPU and fails on the first buffer added without its free."""
    import glob
    import re
    src_dir = os.path.join(os.path.dirname(__file__), "..", "qtip4pf_water_amd", "csrc")
    text = ""
    for path in sorted(glob.glob(os.path.join(src_dir, "*.hpp")) +
                       glob.glob(os.path.join(src_dir, "*.hip"))):
        with open(path) as f:
            text += f.read()
    allocated = set(re.findall(r"dalloc\(ctx,\s*&ctx->(\w+)", text))
    body = text[text.index("void qtp_destroy(qtp_ctx *ctx)"):]
    ptrs = body[body.index("void *ptrs[] = {"):]
    freed = set(re.findall(r"ctx->(\w+)", ptrs[:ptrs.index("};")]))
    assert allocated, "no dalloc calls found"
    assert not (allocated - freed), sorted(allocated - freed)
