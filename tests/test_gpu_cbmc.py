"""Multi-box CBMC trial sets: the one-launch plan over many boxes agrees with
the single-box trial kernel and with the oracle, a captured hipGraph replays
to the same numbers, the autotuner keeps results unchanged, and the lock-step
GCMC driver runs in both precisions."""
import numpy as np
import pytest

from cassandra_amd import cbmc
from cassandra_amd import engine as E
from cassandra_amd import systems as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def driver():
    d = cbmc.MultiBoxCBMC(nboxes=6, seed=11, precision=E.FP64, graph=False, autotune=False)
    yield d
    d.close()


def test_plan_matches_single_box_kernel_and_oracle(driver, oracle_mod):
    print("[gpu] cbmc plan vs single box", flush=True)
    xyz, trials = driver._random_first_fragments()
    trials = trials[:, :, [1, 0, 2], :]
    nrg, ov = driver.plan_a.run(trials.reshape(-1), graph=False)
    nrg = nrg.reshape(driver.nb, -1)
    for i, box in enumerate(driver.boxes):
        single, ov1 = box.trial_energies(cbmc.SP_FRAG_END, trials[i], cbmc=True)
        assert np.array_equal(ov1, ov.reshape(driver.nb, -1)[i])
        ok = ~ov1
        assert np.allclose(nrg[i][ok], single[ok], rtol=1e-12, atol=1e-9)
    # oracle: box 0, CBMC cutoff energies of the same placements
    sysb = driver.boxes_in[0][0]
    b0 = driver.boxes[0]
    live = b0.live.astype(bool)
    xyz0 = b0.download().reshape(-1, 5, 3)[live]
    s2 = S.System(**{**sysb.__dict__})
    s2.nmols = [int(live.sum()), 0, 0, 0, 0]
    s2.coords = [xyz0] + [np.zeros((0, sp.natoms, 3)) for sp in sysb.species[1:]]
    ob = oracle_mod.OracleBox(s2, fold=False)
    out, ovo = ob.widom_energies(cbmc.SP_FRAG_END, trials[0])
    ok = ~ovo[:, 1]
    assert np.array_equal(ovo[:, 1], ov.reshape(driver.nb, -1)[0])
    assert np.max(np.abs(out[ok, 5] - nrg[0][ok]) / np.maximum(np.abs(out[ok, 5]), 1.0)) < 1e-10


def test_graph_replay_and_autotune(driver):
    print("[gpu] cbmc graph replay", flush=True)
    xyz, trials = driver._random_first_fragments()
    x = trials[:, :, [1, 0, 2], :].reshape(-1)
    ref, ovr = driver.plan_a.run(x, graph=False)
    for _ in range(4):
        got, ov = driver.plan_a.run(x, graph=True)
        assert np.array_equal(got, ref) and np.array_equal(ov, ovr)
        driver.boxes[0].compute_system_total_energy()     # eager work between replays
    blk = driver.plan_a.autotune(x)
    assert blk in (64, 128, 256)
    got, ov = driver.plan_a.run(x, graph=True)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-9) and np.array_equal(ov, ovr)


@pytest.mark.parametrize("prec", [E.FP64, E.MIXED])
def test_gcmc_steps_run(prec):
    print(f"[gpu] cbmc gcmc steps prec={prec}", flush=True)
    d = cbmc.MultiBoxCBMC(nboxes=8, seed=5, precision=prec, graph=True, autotune=True)
    n0 = [int(b.live.sum()) for b in d.boxes]
    for _ in range(10):
        d.step()
    n1 = [int(b.live.sum()) for b in d.boxes]
    st = d.stats
    assert st["ins_try"] + st["del_try"] == 80
    assert st["trial_energies"] == 80 * (12 + 48 + 1)
    for b in d.boxes:
        e = b.compute_system_total_energy()
        assert np.isfinite(e["total"])
    print(f"[gpu] molecules {sum(n0)} -> {sum(n1)}, stats {st}", flush=True)
    d.close()
