"""Configurational-bias GCMC of a flexible chain across many boxes at once.

The host side of the CBMC trial loops of Build_Molecule and
Fragment_Placement (fragment_growth.f90:72-707, 1449-2132) for a linear
united-atom chain, run in lock-step over many independent boxes: every
growth stage of every box is one multi-box trial-plan launch on the GPU
(cass_trial_plan_*, optionally replayed from a captured hipGraph).

Per GCMC move and box (insertion or deletion, 50/50):
  * the first fragment (3 sites, a library conformation in a random
    orientation) is placed at kappa_ins random centres of mass; each trial is
    scored at the CBMC cutoff (Compute_Molecule_Nonbond_Inter_Energy with
    cbmc_flag, fragment_growth.f90:519-526);
  * every further fragment adds one site at its library bond angle and
    kappa_dih dihedral angles theta0 + 2 pi k / kappa_dih about the shared
    bond (fragment_growth.f90:1812-1914); each trial is the new site's
    CBMC-cutoff inter energy plus its intramolecular nonbonded energy plus
    the dihedral energy change (fragment_growth.f90:2014-2020);
  * a trial is chosen with probability w_k / W (w = exp(-beta u), zero on
    overlap or beta u >= 708), and the Rosenbluth factor is prod_f W_f / kappa_f;
  * the grown molecule's energy at the full cutoff and the tail-correction
    change enter the acceptance with the CBMC energy removed;
  * a deletion regrows the existing molecule with its own position as trial 1
    (del_flag, fragment_growth.f90:386-390).
Acceptance (ideal-gas reference of the library's bond-angle distribution):
  insert: min(1, exp(beta mu') V / (N + 1) * R * exp(-beta (U_full - U_cbmc)))
  delete: min(1, N / (V exp(beta mu')) / R * exp(+beta (U_full - U_cbmc)))
with mu' the chemical potential less the de Broglie term.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import engine as E
from . import systems as S

KB = S.KBOLTZ


# ------------------------------------------------------------------ model
@dataclass
class ChainModel:
    """United-atom n-pentane (TraPPE-UA): CH3-CH2-CH2-CH2-CH3."""
    bond: float = 1.54
    theta0_deg: float = 114.0
    k_theta_K: float = 31250.0           # K/rad^2 in k (theta - theta0)^2
    opls_K: tuple = (0.0, 355.03, -68.19, 791.32)   # a0..a3 in K
    natoms: int = 5

    def dihedral_energy(self, cos_phi):
        """OPLS a0 + a1 (1 + cos) + a2 (1 - cos 2phi) + a3 (1 + cos 3phi) in
        internal units; Cassandra converts it to the equal RB polynomial
        (input_routines.f90:2104-2109)."""
        a0, a1, a2, a3 = (KB * v for v in self.opls_K)
        c = cos_phi
        return (a0 + a1 + 2 * a2 + a3) + (a1 - 3 * a3) * c - 2 * a2 * c * c + 4 * a3 * c * c * c


def sub_species():
    """Species of the trial sets: the full chain, the two kinds of first
    fragment (end: CH3-CH2-CH2, middle: CH2-CH2-CH2) and single sites."""
    ch3 = dict(mass=15.035, charge=0.0, eps_K=98.0, sigma=3.75)
    ch2 = dict(mass=14.027, charge=0.0, eps_K=46.0, sigma=3.95)
    A3 = S.Atom("CH3", "C", **ch3)
    A2 = S.Atom("CH2", "C", **ch2)
    z = np.zeros((3, 3))
    return [S.species_pentane(),
            S.Species("frag_end", [A3, A2, A2], z),
            S.Species("frag_mid", [A2, A2, A2], z),
            S.Species("site_ch3", [A3], np.zeros((1, 3))),
            S.Species("site_ch2", [A2], np.zeros((1, 3)))]


SP_CHAIN, SP_FRAG_END, SP_FRAG_MID, SP_CH3, SP_CH2 = range(5)
SITE_SPECIES = {0: SP_CH3, 1: SP_CH2, 2: SP_CH2, 3: SP_CH2, 4: SP_CH3}


def fragment_library(model: ChainModel, temperature, n=100, seed=0):
    """Bond angles of a three-site fragment sampled from sin(theta)
    exp(-beta k (theta - theta0)^2) at the box temperature (the distribution
    the reference's fragment libraries hold, Scripts/Frag_Library_Setup)."""
    rng = np.random.default_rng(seed)
    beta = 1.0 / (KB * temperature)
    k = KB * model.k_theta_K
    t0 = math.radians(model.theta0_deg)
    out = []
    while len(out) < n:
        th = rng.uniform(0.0, math.pi)
        acc = math.sin(th) * math.exp(-beta * k * (th - t0) ** 2)
        if rng.random() < acc:
            out.append(th)
    return np.array(out)


def random_rotations(rng, n):
    u1, u2, u3 = rng.random((3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], -1)
    x, y, z, w = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)
    return R.reshape(n, 3, 3)


def place_next(a, b, c, bond, theta, phi):
    """Site d bonded to c with angle b-c-d = theta and dihedral a-b-c-d = phi
    (natural extension reference frame).  Arrays broadcast over leading axes."""
    bc = c - b
    bc /= np.linalg.norm(bc, axis=-1, keepdims=True)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    m = np.cross(n, bc)
    theta, phi = np.broadcast_arrays(np.asarray(theta, dtype=float), np.asarray(phi, dtype=float))
    d2 = np.stack([-bond * np.cos(theta), bond * np.sin(theta) * np.cos(phi), bond * np.sin(theta) * np.sin(phi)], -1)
    return c + d2[..., 0:1] * bc + d2[..., 1:2] * m + d2[..., 2:3] * n


def cos_dihedral(a, b, c, d):
    b1, b2, b3 = b - a, c - b, d - c
    n1 = np.cross(b1, b2)
    n2 = np.cross(b2, b3)
    return (n1 * n2).sum(-1) / (np.linalg.norm(n1, axis=-1) * np.linalg.norm(n2, axis=-1))


# ------------------------------------------------------------- the boxes
def make_boxes(nboxes=48, seed=2024, rc=14.0, density_range=(0.15, 0.6)):
    """Cubic boxes of 28-50 A (rc = 14 A needs L >= 28 A) holding pentane at
    a fraction of the liquid number density (0.62 g/cm^3), per-box T in
    300-450 K and a chemical potential near the ideal-gas value of that
    state, all from one seed."""
    rng = np.random.default_rng(seed)
    species = sub_species()
    rho_liq = 0.62 / 72.15 * 0.6022   # molecules / A^3
    boxes = []
    for b in range(nboxes):
        L = rng.uniform(28.0, 50.0)
        n = max(10, int(rng.uniform(*density_range) * rho_liq * L ** 3))
        T = rng.uniform(300.0, 450.0)
        coords = S.place_random_grid([(species[0], n)], [L] * 3, int(rng.integers(1 << 30)), 2.2, 3.0)
        sys_b = S.System(species, [n, 0, 0, 0, 0], coords + [np.zeros((0, sp.natoms, 3)) for sp in species[1:]],
                         np.array([L] * 3), T, "lj", "cut_tail", rc, charge_style="none", rcut_low=1.0,
                         rcut_cbmc=6.5, name=f"pentane_box{b}")
        mu = KB * T * math.log(n / L ** 3) + rng.uniform(-2.0, 2.0) * KB * T
        boxes.append((sys_b.parsed(), mu))
    return boxes


class MultiBoxCBMC:
    """Lock-step CBMC GCMC in many boxes on one engine."""

    def __init__(self, nboxes=48, kappa_ins=12, kappa_dih=24, seed=2024, precision=E.MIXED, graph=True,
                 margin=64, autotune=True):
        self.model = ChainModel()
        self.kins, self.kdih = kappa_ins, kappa_dih
        self.rng = np.random.default_rng(seed + 1)
        self.graph = graph
        self.boxes_in = make_boxes(nboxes, seed)
        first = self.boxes_in[0][0]
        self.eng = E.Engine(first, max_boxes=nboxes, precision=precision)
        self.boxes, self.mu, self.lib = [], [], []
        for k, (sysb, mu) in enumerate(self.boxes_in):
            cap = sysb.nmols[0] + margin
            box = self.eng.add_box(sysb, max_mols=cap)
            # slots: the live molecules, then dead slots for insertions
            xyz = box.download()
            spec = np.zeros(cap, dtype=np.int32)
            live = np.zeros(cap, dtype=np.int32)
            live[:sysb.nmols[0]] = 1
            full = np.zeros(3 * 5 * cap)
            full[:xyz.size] = xyz
            box.upload(spec, live, full)
            box.xyz = full.reshape(cap, 5, 3)
            self.boxes.append(box)
            self.mu.append(mu)
            self.lib.append(fragment_library(self.model, sysb.temperature, 100, seed + k))
        self.nb = len(self.boxes)
        self.stats = dict(ins_try=0, ins_acc=0, del_try=0, del_acc=0, trial_energies=0)
        # plans: stage A (first fragment), B/C (one new site), D (whole chain, full cutoff)
        ids = [b.id for b in self.boxes]
        self.plan_a = self.eng.trial_plan(ids, [SP_FRAG_END] * self.nb, [self.kins] * self.nb, cbmc=True)
        self.plan_b = self.eng.trial_plan(ids, [SP_CH2] * self.nb, [self.kdih] * self.nb, cbmc=True)
        self.plan_d = self.eng.trial_plan(ids, [SP_CHAIN] * self.nb, [1] * self.nb, cbmc=False)
        self.tile = None
        if autotune:
            xyz = self._random_first_fragments()[0]
            self.tile = self.plan_a.autotune(xyz)
        self.lrc_cache = [self._lrc(b, int(b.live.sum())) for b in self.boxes]

    # -------------------------------------------------------------- helpers
    def _lrc(self, box, n):
        nt = self.eng.topo.ntypes
        nb = np.zeros(nt)
        nb[self.eng.topo.atom_type[0]] += 2 * n     # CH3
        nb[self.eng.topo.atom_type[1]] += 3 * n     # CH2
        return box.lrc(nb)

    def _random_first_fragments(self):
        """kappa_ins placements of a library first fragment per box."""
        nb, k = self.nb, self.kins
        th = np.array([lib[self.rng.integers(len(lib))] for lib in self.lib])
        b = self.model.bond
        frag = np.zeros((nb, 3, 3))           # branch site at the origin, two neighbours
        frag[:, 1, 0] = b
        frag[:, 2, 0] = b * np.cos(th)
        frag[:, 2, 1] = b * np.sin(th)
        frag -= frag.mean(1, keepdims=True)
        R = random_rotations(self.rng, nb)
        frag = np.einsum("bij,baj->bai", R, frag)
        L = np.array([bx.system.box[0] for bx in self.boxes])
        com = (self.rng.random((nb, k, 3)) - 0.5) * L[:, None, None]
        trials = frag[:, None, :, :] + com[:, :, None, :]
        # site order of the end fragment species: CH3, CH2, CH2 -> (neighbour, branch, neighbour)
        order = trials[:, :, [1, 0, 2], :]
        return order.reshape(-1), trials

    def _select(self, u, beta, over):
        """Rosenbluth selection per box: u (nb, k) energies; returns index,
        log W, selected energy and an all-overlap flag."""
        bu = beta[:, None] * u
        w = np.where(over | (bu >= S.MAX_KBT), 0.0, np.exp(-np.minimum(bu, S.MAX_KBT)))
        W = w.sum(1)
        dead = W <= 0.0
        p = np.where(dead[:, None], 1.0 / w.shape[1], w / np.where(dead, 1.0, W)[:, None])
        cum = np.cumsum(p, 1)
        r = self.rng.random(self.nb)
        idx = np.minimum((cum < r[:, None]).sum(1), w.shape[1] - 1)
        return idx, np.log(np.where(dead, 1.0, W)), u[np.arange(self.nb), idx], dead

    # ----------------------------------------------------------------- step
    def step(self):
        """One CBMC insertion or deletion attempt in every box."""
        nb, kin, kd = self.nb, self.kins, self.kdih
        beta = np.array([b.params.beta for b in self.boxes])
        insert = self.rng.random(nb) < 0.5
        nlive = np.array([int(b.live.sum()) for b in self.boxes])
        insert &= nlive < np.array([len(b.live) for b in self.boxes])
        victim = np.full(nb, -1)
        for i in np.nonzero(~insert)[0]:
            alive = np.nonzero(self.boxes[i].live)[0]
            if len(alive) == 0:
                insert[i] = True
                continue
            victim[i] = self.rng.choice(alive)
        # stage A: first fragment = sites 0, 1, 2 grown from the CH3 end
        xyz_a, trials = self._random_first_fragments()
        trials = trials[:, :, [1, 0, 2], :].copy()            # sites 0, 1, 2
        for i in np.nonzero(~insert)[0]:
            trials[i, 0] = self.boxes[i].xyz[victim[i], 0:3]   # trial 1 = the existing molecule
        u_a, ov_a = self._run(self.plan_a, trials.reshape(-1), exclude=victim)
        u_a, ov_a = u_a.reshape(nb, kin), ov_a.reshape(nb, kin)
        ia, logw, uc, dead = self._select(u_a, beta, ov_a)
        ia = np.where(insert, ia, 0)
        uc = u_a[np.arange(nb), ia]
        logR = logw - math.log(kin)
        ucbmc = uc.copy()
        chain = np.zeros((nb, 5, 3))
        chain[:, 0:3] = trials[np.arange(nb), ia]
        e_dih_tot = np.zeros(nb)      # dihedral energy of the grown chain
        e_intra_tot = np.zeros(nb)    # intramolecular nonbonded energy of the grown chain
        # stages B, C: sites 3 and 4 by dihedral trials about the growing bond
        for site in (3, 4):
            th = np.array([lib[self.rng.integers(len(lib))] for lib in self.lib])
            phi0 = 2 * math.pi * self.rng.random(nb)
            phis = phi0[:, None] + 2 * math.pi * np.arange(kd)[None, :] / kd
            a, b_, c = chain[:, site - 3], chain[:, site - 2], chain[:, site - 1]
            new = place_next(a[:, None], b_[:, None], c[:, None], self.model.bond, th[:, None], phis)
            for i in np.nonzero(~insert)[0]:
                new[i, 0] = self.boxes[i].xyz[victim[i], site]
            sp = SITE_SPECIES[site]
            u_b, ov_b = self._run(self.plan_b, new.reshape(-1), exclude=victim, species=sp)
            u_b, ov_b = u_b.reshape(nb, kd), ov_b.reshape(nb, kd)
            cphi = cos_dihedral(a[:, None], b_[:, None], c[:, None], new)
            e_dih = self.model.dihedral_energy(cphi)
            u_intra = self._intra_new(chain, new, site)
            # trial energy = dihedral-energy increment + intramolecular + inter
            # (fragment_growth.f90:2014-2020); the new site adds one dihedral
            u = u_b + u_intra + e_dih
            k, logw, _, dead_b = self._select(u, beta, ov_b | ~np.isfinite(u_intra))
            k = np.where(insert, k, 0)
            sel = np.arange(nb)
            chain[:, site] = new[sel, k]
            ucbmc += u[sel, k]
            logR += logw - math.log(kd)
            dead |= dead_b
            e_dih_tot += e_dih[sel, k]
            e_intra_tot += u_intra[sel, k]
        # stage D: full-cutoff energy of the whole chain
        u_full, ov_full = self._run(self.plan_d, chain.reshape(-1), exclude=victim, species=SP_CHAIN)
        self.stats["trial_energies"] += nb * (kin + 2 * kd + 1)
        for i, box in enumerate(self.boxes):
            n = nlive[i]
            V = box.params.volume
            mu = self.mu[i]
            bi = beta[i]
            if insert[i]:
                self.stats["ins_try"] += 1
                if dead[i] or ov_full[i]:
                    continue
                lrc_new = self._lrc(box, n + 1)
                d_lrc = lrc_new - self.lrc_cache[i]
                ln_acc = bi * mu + math.log(V / (n + 1)) + logR[i] - bi * (u_full[i] + d_lrc + e_dih_tot[i]
                                                                            + e_intra_tot[i] - ucbmc[i])
                if ln_acc >= 0 or self.rng.random() < math.exp(ln_acc):
                    slot = int(np.nonzero(box.live == 0)[0][0])
                    box.live[slot] = 1
                    box.xyz[slot] = chain[i]
                    box.mol_update(slot, chain[i].reshape(-1), live=True)
                    self.lrc_cache[i] = lrc_new
                    self.stats["ins_acc"] += 1
            else:
                self.stats["del_try"] += 1
                lrc_new = self._lrc(box, n - 1)
                d_lrc = self.lrc_cache[i] - lrc_new
                ln_acc = -bi * mu + math.log(n / V) - logR[i] + bi * (u_full[i] + d_lrc + e_dih_tot[i]
                                                                       + e_intra_tot[i] - ucbmc[i])
                if ln_acc >= 0 or self.rng.random() < math.exp(ln_acc):
                    j = int(victim[i])
                    box.live[j] = 0
                    box.mol_update(j, box.xyz[j].reshape(-1), live=False)
                    self.lrc_cache[i] = lrc_new
                    self.stats["del_acc"] += 1

    def _run(self, plan, xyz, exclude, species=None):
        if species is not None or exclude is not None:
            plan.update(species=None if species is None else [species] * self.nb, exclude=exclude)
        return plan.run(xyz, graph=self.graph)

    def _intra_new(self, chain, new, site):
        """Intramolecular nonbonded energy of the new site with the grown
        sites beyond 1-4 (only 1-5 pairs exist in pentane: site 4 with 0)."""
        out = np.zeros(new.shape[:2])
        if site == 4:
            eps, sig = self._lj_pair(SP_CH3, SP_CH3)
            d = new - chain[:, None, 0]
            r2 = (d * d).sum(-1)
            s6 = (sig * sig / r2) ** 3
            u = 4 * eps * (s6 * s6 - s6)
            out = np.where(r2 <= self.boxes[0].params.rcut_cbmcsq, u, 0.0)
        return out

    def _lj_pair(self, a, b):
        t = self.eng.topo
        nt = t.ntypes
        ta = t.atom_type[t.species_atom_offset[a]]
        tb = t.atom_type[t.species_atom_offset[b]]
        return t.eps[ta * nt + tb], t.sig[ta * nt + tb]

    def close(self):
        for p in (self.plan_a, self.plan_b, self.plan_d):
            p.close()
        self.eng.close()
