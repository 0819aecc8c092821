// adjoint.hip -- derivatives of a box's nonbonded energy with respect to the
// per-type LJ parameters, the per-site-kind charges and the coordinates.
//
// No counterpart exists in the reference; the energy being differentiated is
// exactly the one Compute_System_Total_Energy assembles
// (energy_routines.f90:2648-2953): intermolecular LJ + real-space q-q (with
// the A&S erfc of :2143-2163, differentiated as written), the intramolecular
// scaled terms and the periodic image part of intramolecular pairs, the
// reciprocal sum, the self energy and the LJ tail correction.  Lorentz-
// Berthelot mixing (create_nonbond_table.f90:370-392) carries the pair
// derivatives to the per-type parameters.
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "kernels.h"

using namespace cass;

namespace {

constexpr int kAdjTypes = 8;     // atom types
constexpr int kAdjKinds = 16;    // species-atom kinds (charges)
constexpr int kAdjSites = 8;     // sites per molecule
constexpr int kAdjNV = 1 + 6 * kAdjSites;  // energy, then per site (d eps, d sig, d q, force)

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// d/dx of the A&S erfc as computed (t = 1/(1+p x), erfc = t*poly(t)*exp(-x^2)).
__host__ __device__ inline void erfc_as_d(double x, double &f, double &df) {
  const double A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027, A5 = 1.061405429,
               P = 0.3275911;
  const double t = 1.0 / (1.0 + P * x);
  const double poly = A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)));
  const double dpoly = A2 + t * (2.0 * A3 + t * (3.0 * A4 + t * 4.0 * A5));
  const double e = exp(-x * x);
  const double tp = t * poly;
  f = tp * e;
  const double dtp_dt = poly + t * dpoly;
  const double dt_dx = -P * t * t;
  df = dtp_dt * dt_dx * e + tp * (-2.0 * x) * e;
}

struct AdjParams {
  int ntypes, q_style, vdw_sum;
  double rc2_vdw, rc2_coul, rc_low2, rc_coul, alpha, alpha_dsf, dsf1, dsf2;
  double eps_t[kAdjTypes], sig_t[kAdjTypes];
};

// LJ pair with mixing: energy, dU/d eps_ij, dU/d sig_ij, dU/dr / r.
__host__ __device__ inline void lj_pair(const AdjParams &a, double r2, double eps, double sig, double &u,
                                        double &de, double &ds, double &dr_over_r) {
  const double s2 = sig * sig / r2, s6 = s2 * s2 * s2, s12 = s6 * s6;
  u = 4.0 * eps * (s12 - s6);
  de = 4.0 * (s12 - s6);
  ds = sig > 0.0 ? 4.0 * eps * (12.0 * s12 - 6.0 * s6) / sig : 0.0;
  dr_over_r = 4.0 * eps * (-12.0 * s12 + 6.0 * s6) / r2;
  if (a.vdw_sum == CASS_SUM_CUT_SHIFT) {
    const double c2 = sig * sig / a.rc2_vdw, c6 = c2 * c2 * c2, c12 = c6 * c6;
    u -= 4.0 * eps * (c12 - c6);
    de -= 4.0 * (c12 - c6);
    if (sig > 0.0) ds -= 4.0 * eps * (12.0 * c12 - 6.0 * c6) / sig;
  }
}

// q-q pair with charges factored out: E = qi qj g(r); returns g and dg/dr / r.
// intra: the scaled minimum-image part plus the periodic part (Ewald).
__host__ __device__ inline void qq_pair(const AdjParams &a, double r2, bool intra, double qscale, double &g,
                                        double &dg_over_r) {
  const double r = sqrt(r2);
  if (a.q_style == CASS_Q_EWALD) {
    double f, df;
    erfc_as_d(a.alpha * r, f, df);
    const double base = intra ? (qscale - 1.0 + f) : f;   // inter: Eij*erfc
    g = kChargeFactor * base / r;
    const double dbase = df * a.alpha;
    dg_over_r = kChargeFactor * (dbase / r - base / r2) / r;
  } else if (a.q_style == CASS_Q_DSF) {
    const double ec = erfc(a.alpha_dsf * r);
    if (intra) {
      const double v = a.dsf2 * (r - a.rc_coul) - a.dsf1 + ec / r - (1.0 - qscale) / r;
      g = kChargeFactor * v;
      const double dv = a.dsf2 - 2.0 * a.alpha_dsf / 1.7724538509055159 * exp(-a.alpha_dsf * a.alpha_dsf * r2) / r -
                        ec / r2 + (1.0 - qscale) / r2;
      dg_over_r = kChargeFactor * dv / r;
    } else {
      const double v = a.dsf2 * (r - a.rc_coul) - a.dsf1 + ec / r;
      g = kChargeFactor * v;
      const double dv = a.dsf2 - 2.0 * a.alpha_dsf / 1.7724538509055159 * exp(-a.alpha_dsf * a.alpha_dsf * r2) / r -
                        ec / r2;
      dg_over_r = kChargeFactor * dv / r;
    }
  } else if (a.q_style == CASS_Q_CUT || a.q_style == CASS_Q_MINIMUM) {
    const double s = intra ? qscale : 1.0;
    g = kChargeFactor * s / r;
    dg_over_r = -kChargeFactor * s / (r2 * r);
  } else {
    g = 0.0;
    dg_over_r = 0.0;
  }
}

__host__ __device__ inline double mixed_eps(const AdjParams &a, int ti, int tj) {
  const double ei = a.eps_t[ti], ej = a.eps_t[tj];
  if (ti == tj) return ei;
  return (ei <= 1e-7 || ej <= 1e-7) ? 0.0 : sqrt(ei * ej);
}
__host__ __device__ inline double mixed_sig(const AdjParams &a, int ti, int tj) {
  const double si = a.sig_t[ti], sj = a.sig_t[tj];
  if (ti == tj) return si;
  return (si <= 1e-7 || sj <= 1e-7) ? 0.0 : (si + sj) * 0.5;
}

// Ordered molecule pairs (i, j != i): block (bx, i) covers j in
// [bx*blockDim, +blockDim).  Each unordered pair is visited twice; the energy
// is halved and every other quantity is attributed to the sites of i only
// (the visit (j, i) supplies j's side): per site of i, dE/d eps of its type,
// dE/d sigma of its type, dE/d q of its kind, and the force.
__global__ void __launch_bounds__(256) k_adjoint_pairs(BoxDev b, TopoDev tp, AdjParams a, double *part) {
  __shared__ double px[kAdjSites], py[kAdjSites], pz[kAdjSites];
  __shared__ double red[kAdjNV][4];
  const int i = blockIdx.y;
  const int si = b.species[i], fi = b.first[i];
  const int ni = tp.sp_natoms[si], oi = tp.sp_atom_off[si];
  if (threadIdx.x < ni) {
    px[threadIdx.x] = b.x[fi + threadIdx.x];
    py[threadIdx.x] = b.y[fi + threadIdx.x];
    pz[threadIdx.x] = b.z[fi + threadIdx.x];
  }
  __syncthreads();
  double v[kAdjNV];
#pragma unroll
  for (int k = 0; k < kAdjNV; ++k) v[k] = 0.0;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (b.live[i] && j < b.nmols && j != i && b.live[j]) {
    const double4 ci = b.comd[i], cj = b.comd[j];
    double dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
    min_image(b, dx, dy, dz);
    const bool in = b.vdw_minimum || !(sqrt(dx * dx + dy * dy + dz * dz) > b.rcut_max + ci.w + cj.w);
    if (in) {
      const int sj = b.species[j], fj = b.first[j];
      const int nj = tp.sp_natoms[sj], oj = tp.sp_atom_off[sj];
#pragma unroll
      for (int ia = 0; ia < kAdjSites; ++ia) {
        if (ia >= ni) break;
        const int ti = tp.atom_type[oi + ia];
        const double qi = tp.atom_charge[oi + ia];
        double de = 0.0, ds = 0.0, dq = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        for (int ja = 0; ja < nj; ++ja) {
          double rx = px[ia] - b.x[fj + ja], ry = py[ia] - b.y[fj + ja], rz = pz[ia] - b.z[fj + ja];
          min_image(b, rx, ry, rz);
          const double r2 = rx * rx + ry * ry + rz * rz;
          const int tj = tp.atom_type[oj + ja];
          const double qj = tp.atom_charge[oj + ja];
          double fr = 0.0;  // -dE/dr / r on site ia
          if (r2 <= a.rc2_vdw && ti >= 0 && tj >= 0) {
            const double eps = mixed_eps(a, ti, tj), sig = mixed_sig(a, ti, tj);
            double u, deij, dsij, dr;
            lj_pair(a, r2, eps, sig, u, deij, dsij, dr);
            v[0] += 0.5 * u;
            if (eps > 0.0) de += deij * eps / (2.0 * a.eps_t[ti]);
            if (sig > 0.0) ds += dsij * 0.5;
            fr -= dr;
          }
          if (r2 <= a.rc2_coul && a.q_style != CASS_Q_NONE) {
            double g, dg;
            qq_pair(a, r2, false, 1.0, g, dg);
            v[0] += 0.5 * qi * qj * g;
            dq += qj * g;
            fr -= qi * qj * dg;
          }
          fx += fr * rx;
          fy += fr * ry;
          fz += fr * rz;
        }
        v[1 + 6 * ia] += de;
        v[2 + 6 * ia] += ds;
        v[3 + 6 * ia] += dq;
        v[4 + 6 * ia] += fx;
        v[5 + 6 * ia] += fy;
        v[6 + 6 * ia] += fz;
      }
    }
  }
  // block reduction of kAdjNV values
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kAdjNV; ++k) {
    const double sv = wave_sum_d(v[k]);
    if (lane == 0) red[k][wid] = sv;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kAdjNV; k += blockDim.x) {
    double sv = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sv += red[k][w];
    part[((size_t)i * gridDim.x + blockIdx.x) * kAdjNV + k] = sv;
  }
}

// Reciprocal-space derivatives per site: force and dE/dq of its kind, from
// the box sums C_k, S_k of the last full Ewald evaluation.
__global__ void k_adjoint_recip(BoxDev b, TopoDev tp, EwaldDev ew, int nsites, const int *site_mol,
                                const int *site_spa, double *fout, double *dqout) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsites) return;
  const int m = site_mol[s];
  double fx = 0, fy = 0, fz = 0, dq = 0;
  if (b.live[m]) {
    const double x = b.x[s], y = b.y[s], z = b.z[s];
    const double q = tp.atom_charge[site_spa[s]];
    for (int k = 0; k < ew.nvecs; ++k) {
      const double th = ew.hx[k] * x + ew.hy[k] * y + ew.hz[k] * z;
      double sv, cv;
      sincos(th, &sv, &cv);
      const double C = ew.cs[k], S = ew.sn[k], cn = ew.cn[k];
      dq += cn * 2.0 * (C * cv + S * sv);
      const double w = 2.0 * q * cn * (S * cv - C * sv);  // dE/dtheta
      fx -= w * ew.hx[k];
      fy -= w * ew.hy[k];
      fz -= w * ew.hz[k];
    }
  }
  fout[3 * s] = fx * kChargeFactor;
  fout[3 * s + 1] = fy * kChargeFactor;
  fout[3 * s + 2] = fz * kChargeFactor;
  dqout[s] = dq * kChargeFactor;
}

}  // namespace

BoxDev cass_box_view(const cass_engine *e, const cass_box_state &b);
TopoDev cass_topo_view(const cass_engine *e);
EwaldDev cass_ewald_view(const cass_box_state &b);

extern "C" int cass_adjoint(cass_engine *e, int32_t box, double *energy, double *de_deps, double *de_dsig,
                            double *de_dq, double *forces) {
  if (!e || box < 0 || box >= (int)e->boxes.size()) return CASS_ERR_ARG;
  cass_box_state &b = e->boxes[box];
  const cass_params &p = b.p;
  if (p.vdw_style == CASS_VDW_MIE) return cass_fail(e, CASS_ERR_ARG, "adjoint supports LJ only");
  if (!(p.vdw_sum_style == CASS_SUM_CUT || p.vdw_sum_style == CASS_SUM_CUT_TAIL ||
        p.vdw_sum_style == CASS_SUM_CUT_SHIFT))
    return cass_fail(e, CASS_ERR_ARG, "adjoint supports cut, cut_tail and cut_shift");
  if (e->ntypes > kAdjTypes || e->natoms_total > kAdjKinds) return cass_fail(e, CASS_ERR_ARG, "adjoint size limits");
  for (int s = 0; s < e->nspecies; ++s)
    if (e->sp_natoms[s] > kAdjSites) return cass_fail(e, CASS_ERR_ARG, "adjoint: molecule too large");
  if (hipSetDevice(e->device) != hipSuccess) return CASS_ERR_HIP;
  AdjParams a;
  std::memset(&a, 0, sizeof(a));
  a.ntypes = e->ntypes;
  a.q_style = p.charge_style;
  a.vdw_sum = p.vdw_sum_style;
  a.rc2_vdw = p.vdw_style == CASS_VDW_NONE ? -1.0 : p.rcut_vdwsq;
  a.rc2_coul = p.rcut_coulsq;
  a.rc_low2 = p.rcut_lowsq;
  a.rc_coul = p.rcut_coul;
  a.alpha = p.alpha_ewald;
  a.alpha_dsf = p.alpha_dsf;
  a.dsf1 = p.dsf_factor1;
  a.dsf2 = p.dsf_factor2;
  for (int t = 0; t < e->ntypes; ++t) {
    a.eps_t[t] = e->eps[t * e->ntypes + t];
    a.sig_t[t] = e->sig[t * e->ntypes + t];
  }
  const int nt = e->ntypes, nk = e->natoms_total;
  // dE[0] energy, dE[1+t] d/d eps_t, dE[1+kAdjTypes+t] d/d sig_t, dE[1+2*kAdjTypes+k] d/d q_k
  std::vector<double> dE(1 + 2 * kAdjTypes + kAdjKinds, 0.0);
  std::vector<double> F(3 * (size_t)b.nsites, 0.0);
  // intermolecular pairs on the device
  int rc;
  if (b.nmols > 0) {
    const int nbx = (b.nmols + 255) / 256;
    const size_t np = (size_t)nbx * b.nmols * kAdjNV;
    double *part = (double *)cass_scratch(e, np * sizeof(double));
    if (!part) return cass_fail(e, CASS_ERR_NOMEM, "scratch");
    BoxDev bv = cass_box_view(e, b);
    TopoDev tv = cass_topo_view(e);
    hipLaunchKernelGGL(k_adjoint_pairs, dim3(nbx, b.nmols), dim3(256), 0, e->stream, bv, tv, a, part);
    if (hipGetLastError() != hipSuccess) return cass_fail(e, CASS_ERR_HIP, "k_adjoint_pairs launch");
    std::vector<double> hp(np);
    if (hipMemcpyAsync(hp.data(), part, np * sizeof(double), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
      return cass_fail(e, CASS_ERR_HIP, "adjoint copy");
    for (int i = 0; i < b.nmols; ++i) {
      const int si = b.mol_species[i], fi = b.mol_first[i], oi = e->sp_atom_off[si];
      for (int bx = 0; bx < nbx; ++bx) {
        const double *v = hp.data() + ((size_t)i * nbx + bx) * kAdjNV;
        dE[0] += v[0];
        for (int ia = 0; ia < e->sp_natoms[si]; ++ia) {
          const int t = e->atom_type[oi + ia];
          if (t >= 0) {
            dE[1 + t] += v[1 + 6 * ia];
            dE[1 + kAdjTypes + t] += v[2 + 6 * ia];
          }
          dE[1 + 2 * kAdjTypes + oi + ia] += v[3 + 6 * ia];
          for (int d = 0; d < 3; ++d) F[3 * (fi + ia) + d] += v[4 + 6 * ia + d];
        }
      }
    }
  }
  // intramolecular terms on the host (scaled minimum-image + periodic part)
  for (int m = 0; m < b.nmols; ++m) {
    if (!b.live[m]) continue;
    const int s = b.mol_species[m], o = e->sp_atom_off[s], n = e->sp_natoms[s], po = e->sp_pair_off[s];
    const double *xyz = b.xyz_h.data() + 3 * (size_t)b.mol_first[m];
    for (int ia = 0; ia < n; ++ia)
      for (int ja = ia + 1; ja < n; ++ja) {
        const double rx = xyz[3 * ia] - xyz[3 * ja], ry = xyz[3 * ia + 1] - xyz[3 * ja + 1],
                     rz = xyz[3 * ia + 2] - xyz[3 * ja + 2];
        const double r2 = rx * rx + ry * ry + rz * rz;
        double fr = 0.0;
        const int ti = e->atom_type[o + ia], tj = e->atom_type[o + ja];
        const double vs = e->vdw_scale[po + ia * n + ja], qs = e->q_scale[po + ia * n + ja];
        if (r2 <= a.rc2_vdw && ti >= 0 && tj >= 0 && vs != 0.0) {
          const double eps = mixed_eps(a, ti, tj) * vs, sig = mixed_sig(a, ti, tj);
          double u, de, ds, dr;
          lj_pair(a, r2, eps, sig, u, de, ds, dr);
          dE[0] += u;
          if (eps > 0.0) {
            dE[1 + ti] += de * vs * (eps / vs) / (2.0 * a.eps_t[ti]);
            dE[1 + tj] += de * vs * (eps / vs) / (2.0 * a.eps_t[tj]);
          }
          if (sig > 0.0) {
            dE[1 + kAdjTypes + ti] += ds * 0.5;
            dE[1 + kAdjTypes + tj] += ds * 0.5;
          }
          fr -= dr;
        }
        if (r2 <= a.rc2_coul && a.q_style != CASS_Q_NONE) {
          const double qi = e->atom_charge[o + ia], qj = e->atom_charge[o + ja];
          double g, dg;
          qq_pair(a, r2, true, qs, g, dg);
          dE[0] += qi * qj * g;
          dE[1 + 2 * kAdjTypes + o + ia] += qj * g;
          dE[1 + 2 * kAdjTypes + o + ja] += qi * g;
          fr -= qi * qj * dg;
        }
        const size_t fa = 3 * ((size_t)b.mol_first[m] + ia), fb = 3 * ((size_t)b.mol_first[m] + ja);
        F[fa] += fr * rx;
        F[fa + 1] += fr * ry;
        F[fa + 2] += fr * rz;
        F[fb] -= fr * rx;
        F[fb + 1] -= fr * ry;
        F[fb + 2] -= fr * rz;
      }
  }
  // reciprocal space and self energy
  if (p.charge_style == CASS_Q_EWALD) {
    if (!b.ewald_ready) {
      int32_t nv;
      if ((rc = cass_ewald_setup(e, box, &nv))) return rc;
    }
    double er;
    if ((rc = cass_ewald_full(e, box, &er))) return rc;
    dE[0] += er;
    std::vector<int32_t> site_mol(b.nsites), site_spa(b.nsites);
    for (int m = 0; m < b.nmols; ++m)
      for (int aa = 0; aa < e->sp_natoms[b.mol_species[m]]; ++aa) {
        site_mol[b.mol_first[m] + aa] = m;
        site_spa[b.mol_first[m] + aa] = e->sp_atom_off[b.mol_species[m]] + aa;
      }
    const size_t ns = b.nsites;
    char *s = (char *)cass_scratch(e, ns * (2 * sizeof(int32_t) + 4 * sizeof(double)) + 64);
    if (!s) return cass_fail(e, CASS_ERR_NOMEM, "scratch");
    int32_t *d_mol = (int32_t *)s, *d_spa = d_mol + ns;
    double *d_f = (double *)(s + ((ns * 2 * sizeof(int32_t) + 15) / 16) * 16);
    double *d_q = d_f + 3 * ns;
    std::vector<double> hf(3 * ns), hq(ns);
    BoxDev bv = cass_box_view(e, b);
    TopoDev tv = cass_topo_view(e);
    EwaldDev ew = cass_ewald_view(b);
    if (hipMemcpyAsync(d_mol, site_mol.data(), ns * 4, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
        hipMemcpyAsync(d_spa, site_spa.data(), ns * 4, hipMemcpyHostToDevice, e->stream) != hipSuccess)
      return cass_fail(e, CASS_ERR_HIP, "adjoint upload");
    hipLaunchKernelGGL(k_adjoint_recip, dim3((ns + 255) / 256), dim3(256), 0, e->stream, bv, tv, ew, (int)ns, d_mol,
                       d_spa, d_f, d_q);
    if (hipGetLastError() != hipSuccess) return cass_fail(e, CASS_ERR_HIP, "k_adjoint_recip launch");
    if (hipMemcpyAsync(hf.data(), d_f, 3 * ns * 8, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipMemcpyAsync(hq.data(), d_q, ns * 8, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
      return cass_fail(e, CASS_ERR_HIP, "adjoint copy");
    for (size_t i = 0; i < 3 * ns; ++i) F[i] += hf[i];
    for (size_t i = 0; i < ns; ++i)
      if (b.live[site_mol[i]]) dE[1 + 2 * kAdjTypes + site_spa[i]] += hq[i];
    // self energy -cf alpha/sqrt(pi) sum q^2
    for (int m = 0; m < b.nmols; ++m) {
      if (!b.live[m]) continue;
      const int sp = b.mol_species[m];
      for (int aa = 0; aa < e->sp_natoms[sp]; ++aa) {
        const double q = e->atom_charge[e->sp_atom_off[sp] + aa];
        dE[0] += -q * q * kChargeFactor * p.alpha_ewald / kRootPi;
        dE[1 + 2 * kAdjTypes + e->sp_atom_off[sp] + aa] += -2.0 * q * kChargeFactor * p.alpha_ewald / kRootPi;
      }
    }
  }
  // LJ tail correction (energy_routines.f90:2956-3023) with mixed parameters
  if (p.vdw_sum_style == CASS_SUM_CUT_TAIL && p.vdw_style == CASS_VDW_LJ) {
    std::vector<double> nb(nt, 0.0);
    for (int m = 0; m < b.nmols; ++m) {
      if (!b.live[m]) continue;
      const int sp = b.mol_species[m];
      for (int aa = 0; aa < e->sp_natoms[sp]; ++aa) {
        const int t = e->atom_type[e->sp_atom_off[sp] + aa];
        if (t >= 0) nb[t] += 1.0;
      }
    }
    const double pre = 2.0 * kPi / p.volume;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        const double eps = mixed_eps(a, i, j), sig = mixed_sig(a, i, j);
        const double s2 = sig * sig, s6 = s2 * s2 * s2, s12 = s6 * s6;
        const double w = pre * nb[i] * nb[j];
        const double f = s12 / (9.0 * p.rcut9) - s6 / (3.0 * p.rcut3);
        dE[0] += w * 4.0 * eps * f;
        if (eps > 0.0) {
          dE[1 + i] += w * 4.0 * f * eps / (2.0 * a.eps_t[i]);
          dE[1 + j] += w * 4.0 * f * eps / (2.0 * a.eps_t[j]);
        }
        if (sig > 0.0) {
          const double dfs = (12.0 * s12 / (9.0 * p.rcut9) - 6.0 * s6 / (3.0 * p.rcut3)) / sig;
          dE[1 + kAdjTypes + i] += w * 4.0 * eps * dfs * 0.5;
          dE[1 + kAdjTypes + j] += w * 4.0 * eps * dfs * 0.5;
        }
      }
  }
  if (energy) *energy = dE[0];
  for (int t = 0; t < nt; ++t) {
    if (de_deps) de_deps[t] = dE[1 + t];
    if (de_dsig) de_dsig[t] = dE[1 + kAdjTypes + t];
  }
  if (de_dq)
    for (int k = 0; k < nk; ++k) de_dq[k] = dE[1 + 2 * kAdjTypes + k];
  if (forces) std::memcpy(forces, F.data(), sizeof(double) * F.size());
  return CASS_OK;
}
