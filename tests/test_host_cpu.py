"""Host-side logic on the CPU: the C ABI library loads and exports every
entry point its header declares, the host functions that need no GPU
(parameter derivation, the reference's RNG and Metropolis test) agree with
the oracle and the reference, and the input-format helpers round-trip."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from cassandra_amd import abi, formats
from cassandra_amd import engine as E
from cassandra_amd import systems as S
from tests.conftest import ROOT, config_by_name, golden_json


def header_symbols():
    with open(os.path.join(ROOT, "include", "cass_gpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cass_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = E.load_library()
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_params_derive_matches_oracle(oracle_mod):
    for name in ("lj500", "spce512", "sty_shift_dsf", "sty_switch_cut", "sty_ortho_ewald", "sty_minimum_minimum"):
        system = config_by_name(name).parsed()
        p = E.derive(system)
        q = oracle_mod.derive(system)
        for f, _ in abi.CassParams._fields_:
            a, b = getattr(p, f), getattr(q, f)
            if isinstance(a, (int, float)):
                assert a == b or (math.isnan(a) and math.isnan(b)), (name, f, a, b)
            else:
                assert list(a) == list(b), (name, f)


def test_host_rng_matches_reference():
    lib = E.load_library()
    for case in golden_json("rng.json"):
        r = E.rng(*case["seeds"])
        u = [lib.cass_rng_uniform(C.byref(r)) for _ in case["draws"]]
        assert u == [float(x) for x in case["draws"]]
        assert [lib.cass_rng_int(C.byref(r)) for _ in case["ints"]] == case["ints"]


def test_accept_or_reject_draws_only_when_needed():
    """accept_or_reject.f90:45-62: no draw for ln_pacc <= 0 or >= max_kBT."""
    lib = E.load_library()
    r = E.rng(11, 13)
    s0 = list(r.s)
    assert lib.cass_accept_or_reject(C.byref(r), -1.0) == 1 and list(r.s) == s0
    assert lib.cass_accept_or_reject(C.byref(r), 708.0) == 0 and list(r.s) == s0
    r2 = E.rng(11, 13)
    u = lib.cass_rng_uniform(C.byref(r2))
    r = E.rng(11, 13)
    acc = lib.cass_accept_or_reject(C.byref(r), 0.5)
    assert acc == (u <= math.exp(-0.5)) and list(r.s) == list(r2.s)


@pytest.mark.parametrize("text,expected", [("3.166", 3.166), ("120.0", 120.0), ("0.00001", 1e-05),
                                           ("-12.5", -12.5), ("7", 7.0)])
def test_string_to_double_simple(text, expected):
    assert formats.string_to_double(text) == expected


def test_string_to_double_reproduces_reference_rounding(oracle_mod):
    """io_utilities.f90:331-429: an exponent multiplies by a default-REAL power
    of ten, so 1e-05 reads as float32(1e-5); the reference's alpha_ewald
    (recorded in the fixtures) follows from exactly that value."""
    assert formats.string_to_double("1e-05") == float(np.float32(1e-5))
    assert formats.string_to_double("-0.8476") == -0.8476000000000001
    g = golden_json("spce512_energy.json")
    system = config_by_name("spce512").parsed()
    assert E.derive(system).alpha_ewald == float(g["ewald"]["alpha"])


def test_writers_round_trip(tmp_path):
    system = config_by_name("sty_cut_ewald")
    formats.write_xyz(system, str(tmp_path / "f.xyz"))
    s2 = S.System(**{**system.__dict__})
    formats.read_xyz(str(tmp_path / "f.xyz"), s2)
    for a, b in zip(system.coords, s2.coords):
        assert np.array_equal(a, b)
    inp = formats.write_case(system, str(tmp_path / "case"), "case")
    text = open(inp).read()
    assert "coul ewald 9.0 0.00001" in text and "read_config 100" in text


def test_topology_tables():
    system = S.with_probe(config_by_name("spce512"), S.species_co2()).parsed()
    t = abi.Topology(system)
    # O_w, H_w, C_co2, O_co2
    assert t.ntypes == 4
    eps = t.eps.reshape(4, 4)
    sig = t.sig.reshape(4, 4)
    assert eps[1, 1] == 0.0 and eps[0, 1] == 0.0          # zero-epsilon H mixes to zero
    assert sig[0, 3] == (3.166 + 3.05) * 0.5 or abs(sig[0, 3] - (system.species[0].atoms[0].sigma
                                                               + system.species[1].atoms[1].sigma) * 0.5) == 0
    vs = t.vdw_intra_scale[:9].reshape(3, 3)
    assert vs[0, 1] == 0.0 and vs[1, 2] == 0.0            # 1-2 and 1-3 of water excluded


def test_cbmc_chain_growth_geometry():
    """The CBMC driver's site placement (natural extension reference frame):
    the new site sits at the bond length from c, at angle theta to b-c and
    at dihedral phi about b-c, for broadcast batches of trials."""
    from cassandra_amd import cbmc
    r = np.random.default_rng(0)
    a, b, c = [r.normal(size=(5, 3)) for _ in range(3)]
    th = np.full(5, 1.9)
    ph = r.uniform(-3.0, 3.0, size=(5, 7))
    d = cbmc.place_next(a[:, None], b[:, None], c[:, None], 1.54, th[:, None], ph)
    assert d.shape == (5, 7, 3)
    assert np.allclose(np.linalg.norm(d - c[:, None], axis=-1), 1.54)
    cb = b[:, None] - c[:, None]
    cosang = (cb * (d - c[:, None])).sum(-1) / (np.linalg.norm(cb, axis=-1) * 1.54)
    assert np.allclose(cosang, math.cos(1.9))
    full = lambda x: np.broadcast_to(x[:, None], d.shape)
    assert np.allclose(cbmc.cos_dihedral(full(a), full(b), full(c), d), np.cos(ph))
