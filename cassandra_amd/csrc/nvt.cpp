// nvt.cpp -- the reference's NVT move loop on top of the engine:
// nvtmc_driver.f90:94-152 choosing between Translate (move_translate.f90:23-360)
// and Rotate (move_rotate.f90:23-520), with the same random-number draws, the
// same energy bookkeeping and accept_or_reject.  Every energy is evaluated on
// the GPU; the host keeps the canonical coordinates and the RNG.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "engine.h"
#include "kernels.h"

using namespace cass;

int cass_mol_pair_trials(cass_engine *e, int32_t box, int32_t mol, int32_t cbmc, int32_t ntrial,
                         const double *xyz2, double *ev, double *eq, int32_t *ov);

extern "C" int cass_nvt_run(cass_engine *e, int32_t box, cass_rng *rng, const cass_nvt_options *opt,
                            int64_t nsteps, cass_energy *en, cass_nvt_record *rec) {
  if (!e || !rng || !opt || !en || box < 0 || box >= (int)e->boxes.size()) return CASS_ERR_ARG;
  cass_box_state &b = e->boxes[box];
  const int sp = opt->species;
  if (sp < 0 || sp >= e->nspecies) return CASS_ERR_ARG;
  const int n = e->sp_natoms[sp];
  bool has_q = false;
  for (int a = 0; a < n; ++a) has_q = has_q || e->atom_charge[e->sp_atom_off[sp] + a] != 0.0;
  const bool charged = b.p.charge_style == CASS_Q_EWALD && has_q;
  if (charged && !b.ewald_ready) return cass_fail(e, CASS_ERR_STATE, "Ewald sums not set up");
  std::vector<int> mols;  // locate order of the mobile species
  for (int m = 0; m < b.nmols; ++m)
    if (b.mol_species[m] == sp && b.live[m]) mols.push_back(m);
  const int nm = (int)mols.size();
  if (nm == 0) return cass_fail(e, CASS_ERR_STATE, "no molecules to move");
  std::vector<double> two(6 * (size_t)n), xn(3 * (size_t)n);
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int64_t step = 0; step < nsteps; ++step) {
    const double rno = cass_rng_uniform(rng);
    const bool rot = !(rno <= opt->prob_translate);
    (void)cass_rng_uniform(rng);  // species pick (one mobile species)
    const int im = (int)(cass_rng_uniform(rng) * nm);
    const int m = mols[im];
    const double *xo = b.xyz_h.data() + 3 * (size_t)b.mol_first[m];
    const double4 co = b.comd_h[m];
    double4 cn = co;
    std::memcpy(xn.data(), xo, sizeof(double) * 3 * n);
    if (!rot) {
      const double dx = (2.0 * cass_rng_uniform(rng) - 1.0) * opt->max_disp;
      const double dy = (2.0 * cass_rng_uniform(rng) - 1.0) * opt->max_disp;
      const double dz = (2.0 * cass_rng_uniform(rng) - 1.0) * opt->max_disp;
      for (int a = 0; a < n; ++a) {
        xn[3 * a] = xn[3 * a] + dx;
        xn[3 * a + 1] = xn[3 * a + 1] + dy;
        xn[3 * a + 2] = xn[3 * a + 2] + dz;
      }
      cn.x = co.x + dx;
      cn.y = co.y + dy;
      cn.z = co.z + dz;
      // Fold_Molecule (minimum_image_separation.f90:162-240), cubic boxes
      if (b.p.box_shape == CASS_BOX_CUBIC) {
        double *c[3] = {&cn.x, &cn.y, &cn.z};
        for (int d = 0; d < 3; ++d) {
          const double L = b.p.cell[4 * d], h = b.p.hlength[d];
          if (*c[d] > h) {
            *c[d] = *c[d] - L;
            for (int a = 0; a < n; ++a) xn[3 * a + d] = xn[3 * a + d] - L;
          } else if (*c[d] < -h) {
            *c[d] = *c[d] + L;
            for (int a = 0; a < n; ++a) xn[3 * a + d] = xn[3 * a + d] + L;
          }
        }
      }
    } else {
      double r[9];
      if (opt->linear) {
        // Rotate_Molecule_Eulerian (move_rotate.f90:468-520)
        const double th = std::acos(1.0 - 2.0 * cass_rng_uniform(rng));
        const double ph = (1.0 - 2.0 * cass_rng_uniform(rng)) * kPi;
        const double ps = (1.0 - 2.0 * cass_rng_uniform(rng)) * kPi;
        r[0] = std::cos(ps) * std::cos(ph) - std::cos(th) * std::sin(ph) * std::sin(ps);
        r[1] = std::cos(ps) * std::sin(ph) + std::cos(th) * std::cos(ph) * std::sin(ps);
        r[2] = std::sin(ps) * std::sin(th);
        r[3] = -std::sin(ps) * std::cos(ph) - std::cos(th) * std::sin(ph) * std::cos(ps);
        r[4] = -std::sin(ps) * std::sin(ph) + std::cos(th) * std::cos(ph) * std::cos(ps);
        r[5] = std::cos(ps) * std::sin(th);
        r[6] = std::sin(th) * std::sin(ph);
        r[7] = -std::sin(th) * std::cos(ph);
        r[8] = std::cos(th);
      } else {
        // Rotate_Molecule_Axis (move_rotate.f90:362-463)
        double p1 = 0, p2 = 0, p3 = 0;
        const int axis = (int)(3.0 * cass_rng_uniform(rng)) + 1;
        const double ang = (2.0 * cass_rng_uniform(rng) - 1.0) * opt->max_rot;
        if (axis == 1) p1 = ang;
        else if (axis == 2) p2 = ang;
        else p3 = ang;
        const double c1 = std::cos(p1), c2 = std::cos(p2), c3 = std::cos(p3);
        const double s1 = std::sin(p1), s2 = std::sin(p2), s3 = std::sin(p3);
        r[0] = c1 * c3 - s1 * c2 * s3;
        r[1] = s1 * c3 + c1 * c2 * s3;
        r[2] = s2 * s3;
        r[3] = -c1 * s3 - s1 * c2 * c3;
        r[4] = -s1 * s3 + c1 * c2 * c3;
        r[5] = s2 * c3;
        r[6] = s1 * s2;
        r[7] = -c1 * s2;
        r[8] = c2;
      }
      for (int a = 0; a < n; ++a) {
        const double px = xo[3 * a] - co.x, py = xo[3 * a + 1] - co.y, pz = xo[3 * a + 2] - co.z;
        xn[3 * a] = (r[0] * px + r[1] * py + r[2] * pz) + co.x;
        xn[3 * a + 1] = (r[3] * px + r[4] * py + r[5] * pz) + co.y;
        xn[3 * a + 2] = (r[6] * px + r[7] * py + r[8] * pz) + co.z;
      }
    }
    // old and new inter energies in one launch
    std::memcpy(two.data(), xo, sizeof(double) * 3 * n);
    std::memcpy(two.data() + 3 * n, xn.data(), sizeof(double) * 3 * n);
    double ev[2], eq[2];
    int32_t ov[2];
    int rc = cass_mol_pair_trials(e, box, m, 0, 2, two.data(), ev, eq, ov);
    if (rc) return rc;
    if (ov[0]) return cass_fail(e, CASS_ERR_STATE, "molecule energy too high before the move");
    bool accept = false;
    double ln_pacc = nan;
    if (!ov[1]) {
      double dE = 0.0, erm = 0.0;
      if (charged) {
        rc = cass_ewald_move(e, box, m, rot ? CASS_MOVE_ROTATION : CASS_MOVE_TRANSLATION, xn.data(), &erm, nullptr);
        if (rc) return rc;
        dE = erm - en->reciprocal;
      }
      if (!rot) dE = dE + (ev[1] - ev[0]) + (eq[1] - eq[0]);
      else dE = dE + ev[1] - ev[0] + eq[1] - eq[0];
      ln_pacc = b.p.beta * dE;
      accept = cass_accept_or_reject(rng, ln_pacc) != 0;
      if (accept) {
        en->total = en->total + dE;
        en->inter = en->inter + dE;
        en->inter_vdw = en->inter_vdw + ev[1] - ev[0];
        en->inter_q = en->inter_q + eq[1] - eq[0];
        if (charged) en->reciprocal = erm;
        double com[3] = {cn.x, cn.y, cn.z};
        if (rot) {  // Get_COM after an accepted rotation (move_rotate.f90:287)
          double tm = 0, cx = 0, cy = 0, cz = 0;
          for (int a = 0; a < n; ++a) {
            const double mass = e->atom_mass[e->sp_atom_off[sp] + a];
            tm = tm + mass;
            cx = cx + mass * xn[3 * a];
            cy = cy + mass * xn[3 * a + 1];
            cz = cz + mass * xn[3 * a + 2];
          }
          com[0] = cx / tm;
          com[1] = cy / tm;
          com[2] = cz / tm;
        }
        const double dcom = co.w;
        if ((rc = cass_mol_update(e, box, m, 1, xn.data(), com, &dcom))) return rc;
        if (charged && (rc = cass_ewald_commit(e, box))) return rc;
      } else if (charged) {
        if ((rc = cass_ewald_revert(e, box))) return rc;
      }
    }
    if (rec) {
      if (rec->kind) rec->kind[step] = rot ? 1 : 0;
      if (rec->accepted) rec->accepted[step] = accept ? 1 : 0;
      if (rec->energy) rec->energy[step] = en->total;
      if (rec->ln_pacc) rec->ln_pacc[step] = ln_pacc;
      if (rec->rng_state) std::memcpy(rec->rng_state + 5 * step, rng->s, sizeof(rng->s));
    }
  }
  return CASS_OK;
}
