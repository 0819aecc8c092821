"""The type-pair core radii of calc_rmin (create_nonbond_table.f90:285-301,
398-417) and the cell-list overlap pre-screen that uses them (check_overlap,
sector_routines.f90:160-233) in the oracle:
  - r_min is where the pair's LJ energy reaches U_max = max(U_max_base -
    min_U_q, 0); the test finds min_U_q as the lowest charge product over
    every pair of atoms of the two types (the reference takes it from the
    extremes of each type's charges);
  - the overlap flags of the reference's own Widom placements against a
    numpy scan of every site pair."""
import numpy as np
import pytest

from cassandra_amd import systems as S
from tests.conftest import config_by_name, golden_npz

CHARGE_FACTOR = 138935.4558


def widom_oracle(oracle_mod, case):
    system = S.with_probe(config_by_name(case), S.species_co2()).parsed()
    return system, oracle_mod.OracleBox(system)


@pytest.mark.parametrize("case", ["lj2000", "spce512", "co2n2"])
def test_rminsq_reaches_u_max(oracle_mod, case):
    system, ob = widom_oracle(oracle_mod, case)
    p, t = ob.params, ob.topo
    u_base = 708.0 / p.beta
    r = ob.rminsq_table(u_base)
    nt = t.ntypes
    eps, sig = t.eps.reshape(nt, nt), t.sig.reshape(nt, nt)
    q_of_type = [t.atom_charge[t.atom_type == k] for k in range(nt)]
    checked = 0
    for i in range(nt):
        for j in range(nt):
            if not eps[i, j] > 0.0:
                assert r[i, j] == p.rcut_lowsq
                continue
            u_q = 0.0
            if system.charge_style != "none":
                prods = np.outer(q_of_type[i], q_of_type[j]).ravel()
                lo = prods.min()
                u_q = lo * CHARGE_FACTOR / (np.sqrt(p.rcut_lowsq) if lo <= 0.0 else sig[i, j])
            u_max = max(u_base - u_q, 0.0)
            s6 = (sig[i, j] ** 2 / r[i, j]) ** 3
            assert abs(4.0 * eps[i, j] * (s6 * s6 - s6) - u_max) <= 1e-9 * max(u_max, 1.0), (i, j)
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("case", ["lj2000", "spce512"])
def test_check_overlap_matches_site_scan(oracle_mod, case):
    system, ob = widom_oracle(oracle_mod, case)
    p, t = ob.params, ob.topo
    sp = len(system.species) - 1
    place = golden_npz(f"{case}_co2_widom.npz")["placements"][:60]
    base, bov = ob.widom_energies(sp, place)
    r = ob.rminsq_table(708.0 / p.beta)
    ob.widom_set_rminsq(r)
    try:
        out, ov = ob.widom_energies(sp, place)
    finally:
        ob.widom_set_rminsq(None)
    L = p.basis_length[0]
    xyz = ob.xyz().reshape(-1, 3)
    site_spa = np.concatenate([t.species_atom_offset[s] + np.arange(t.species_natoms[s]) for s in ob.spec])
    tj = t.atom_type[site_spa]
    oi, ni = t.species_atom_offset[sp], t.species_natoms[sp]
    ti = t.atom_type[oi:oi + ni]
    thr = np.where((ti[:, None] >= 0) & (tj[None, :] >= 0), r[np.maximum(tj, 0)[None, :], np.maximum(ti, 0)[:, None]],
                   p.rcut_lowsq)
    scan = np.zeros(len(place), bool)
    for k, pl in enumerate(place.reshape(len(place), ni, 3)):
        d = xyz[None, :, :] - pl[:, None, :]
        d -= L * np.rint(d / L)
        scan[k] = np.any((d * d).sum(-1) < thr)
    assert np.array_equal(ov[:, 0], bov[:, 0] | scan)
    assert np.array_equal(ov[:, 1], bov[:, 1] | scan)
    # the pre-screen rejects more than the rcut_low test alone
    assert scan.sum() > bov[:, 0].sum()
