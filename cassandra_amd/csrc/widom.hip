// widom.hip -- batched Widom test insertions of a rigid molecule into a frame
// (the insertion loop of Widom_Insert, widom_insert.f90:319-489).
//
// One lane scores one insertion: the inter energy at the full and at the
// CBMC cutoff (overlap semantics of Build_Molecule, fragment_growth.f90:519-544),
// the reciprocal-space change, and widom_var = exp(-beta dE).
//
// Insertions are scored in spatial order: a first pass bins the insertion
// centres into a cell grid and sorts the insertion indices by cell, so the 64
// lanes of a wavefront hold insertions from one small region of the box.
// The frame loop is then uniform across the wavefront: a frame molecule that
// no lane reaches is skipped by the whole wave at the cost of one centre
// distance test, and the pair code runs only where some lane is in range.
// Results are written by insertion index, so the sort changes no value.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace cass {

constexpr int kWBlock = 256;

// Frame molecules as the insertion kernel reads them: live molecules in index
// order, their sites packed (x, y, z, charge) with the atom type beside them.
struct FrameMol {
  double4 com;  // (xcom, ycom, zcom, max_dcom)
  int first, n, pad0, pad1;
};

struct WidomFrame {
  const FrameMol *__restrict__ mol;
  const double4 *__restrict__ site;
  const int *__restrict__ type;
  int nmols;
};

// Per probe site a and frame atom type t: mixed (eps, sig, mie n, mie m),
// row-major [a][t].
struct WidomPairTab {
  const double4 *__restrict__ p;
  int ntypes;
};

// Centre of insertion i in fractional coordinates of the box, in [0, 1).
__device__ inline void widom_centre_frac(const WidomArgs &w, const BoxDev &b, int64_t i, double &fx, double &fy,
                                         double &fz) {
  if (w.place) {
    double tm = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int a = 0; a < w.natoms; ++a) {
      const double m = w.mass[a];
      tm += m;
      cx += m * w.place[(i * w.natoms + a) * 3];
      cy += m * w.place[(i * w.natoms + a) * 3 + 1];
      cz += m * w.place[(i * w.natoms + a) * 3 + 2];
    }
    if (tm > 0.0) {
      cx /= tm;
      cy /= tm;
      cz /= tm;
    }
    fx = b.Li[0] * cx + b.Li[3] * cy + b.Li[6] * cz;
    fy = b.Li[1] * cx + b.Li[4] * cy + b.Li[7] * cz;
    fz = b.Li[2] * cx + b.Li[5] * cy + b.Li[8] * cz;
    fx -= floor(fx + 0.5);
    fy -= floor(fy + 0.5);
    fz -= floor(fz + 0.5);
    fx += 0.5;
    fy += 0.5;
    fz += 0.5;
  } else {
    // the centre draws of place_probe
    fx = counter_uniform(w.seed, (uint64_t)i, 0);
    fy = counter_uniform(w.seed, (uint64_t)i, 1);
    fz = counter_uniform(w.seed, (uint64_t)i, 2);
  }
}

__global__ void k_widom_keys(WidomArgs w, BoxDev b, int g0, int g1, int g2, unsigned *keys, int *vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.n_ins) return;
  double fx, fy, fz;
  widom_centre_frac(w, b, i, fx, fy, fz);
  const int cx = min(g0 - 1, max(0, (int)(fx * g0)));
  const int cy = min(g1 - 1, max(0, (int)(fy * g1)));
  const int cz = min(g2 - 1, max(0, (int)(fz * g2)));
  keys[i] = (unsigned)((cz * g1 + cy) * g0 + cx);
  vals[i] = (int)i;
}

// place_probe with the site count fixed at compile time.
template <int NS>
__device__ inline void place_probe_n(const WidomArgs &w, const BoxDev &b, uint64_t i, double (&px)[NS],
                                     double (&py)[NS], double (&pz)[NS]) {
  const double u0 = counter_uniform(w.seed, i, 0), u1 = counter_uniform(w.seed, i, 1);
  const double u2 = counter_uniform(w.seed, i, 2), u3 = counter_uniform(w.seed, i, 3);
  const double u4 = counter_uniform(w.seed, i, 4), u5 = counter_uniform(w.seed, i, 5);
  const double fx = u0 - 0.5, fy = u1 - 0.5, fz = u2 - 0.5;
  const double cx = b.Lc[0] * fx + b.Lc[3] * fy + b.Lc[6] * fz;
  const double cy = b.Lc[1] * fx + b.Lc[4] * fy + b.Lc[7] * fz;
  const double cz = b.Lc[2] * fx + b.Lc[5] * fy + b.Lc[8] * fz;
  const double twopi = 6.283185307179586;
  const double r1 = sqrt(1.0 - u3), r2 = sqrt(u3);
  double s4, c4, s5, c5;
  sincos(twopi * u4, &s4, &c4);
  sincos(twopi * u5, &s5, &c5);
  const double qx = r1 * s4, qy = r1 * c4;
  const double qz = r2 * s5, qw = r2 * c5;
  const double R00 = 1 - 2 * (qy * qy + qz * qz), R01 = 2 * (qx * qy - qz * qw), R02 = 2 * (qx * qz + qy * qw);
  const double R10 = 2 * (qx * qy + qz * qw), R11 = 1 - 2 * (qx * qx + qz * qz), R12 = 2 * (qy * qz - qx * qw);
  const double R20 = 2 * (qx * qz - qy * qw), R21 = 2 * (qy * qz + qx * qw), R22 = 1 - 2 * (qx * qx + qy * qy);
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    const double rx = w.ref[3 * a], ry = w.ref[3 * a + 1], rz = w.ref[3 * a + 2];
    px[a] = R00 * rx + R01 * ry + R02 * rz + cx;
    py[a] = R10 * rx + R11 * ry + R12 * rz + cy;
    pz[a] = R20 * rx + R21 * ry + R22 * rz + cz;
  }
}

// Reciprocal-space energy change of adding the test molecule:
// charge_factor * sum_k Cn (2 C dC + dC^2 + 2 S dS + dS^2) with
// exp(i h.r) = Ea^nx Eb^ny Ec^nz built by phase recurrence along the k rows
// (h.r = 2 pi n.s, s the fractional coordinates, for the k sets of both
// branches of Ewald_Reciprocal_Lattice_Vector_Setup).  FP64 throughout.
template <int NS>
__device__ inline double recip_delta_n(const EwaldDev &ew, const BoxDev &b, const WidomArgs &w,
                                       const double (&px)[NS], const double (&py)[NS], const double (&pz)[NS]) {
  cplx Ea[NS], Eb[NS], Ec[NS], Pb[NS], Pc[NS];
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    const double sx = b.Li[0] * px[a] + b.Li[3] * py[a] + b.Li[6] * pz[a];
    const double sy = b.Li[1] * px[a] + b.Li[4] * py[a] + b.Li[7] * pz[a];
    const double sz = b.Li[2] * px[a] + b.Li[5] * py[a] + b.Li[8] * pz[a];
    double s, c;
    sincos(kTwoPi * sx, &s, &c);
    Ea[a] = {c, s};
    sincos(kTwoPi * sy, &s, &c);
    Eb[a] = {c, s};
    sincos(kTwoPi * sz, &s, &c);
    Ec[a] = {c, s};
    Pb[a] = {1.0, 0.0};
    Pc[a] = {1.0, 0.0};
  }
  int cur_ny = -1000000, cur_nz = -1000000;
  double de = 0.0;
  for (int r = 0; r < ew.nrows; ++r) {
    const int4 row = ew.rows[r];
    const int ny = row.x, nz = row.y, nx0 = row.z, cnt = row.w;
    if (nz != cur_nz) {
#pragma unroll
      for (int a = 0; a < NS; ++a) Pc[a] = cpow_int(Ec[a], nz);
      cur_nz = nz;
      cur_ny = -1000000;
    }
    if (ny == cur_ny + 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) Pb[a] = cmul(Pb[a], Eb[a]);
    } else {
#pragma unroll
      for (int a = 0; a < NS; ++a) Pb[a] = cpow_int(Eb[a], ny);
    }
    cur_ny = ny;
    cplx T[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      const cplx base = cmul(Pb[a], Pc[a]);
      T[a] = nx0 == 0 ? base : cmul(base, cpow_int(Ea[a], nx0));
    }
    const int k0 = ew.row_k0[r];
    for (int kk = 0; kk < cnt; ++kk) {
      double dc = 0.0, ds = 0.0;
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        dc += w.q[a] * T[a].re;
        ds += w.q[a] * T[a].im;
        T[a] = cmul(T[a], Ea[a]);
      }
      const int k = k0 + kk;
      const double c0 = ew.cs[k], s0 = ew.sn[k];
      de += ew.cn[k] * (2.0 * c0 * dc + dc * dc + 2.0 * s0 * ds + ds * ds);
    }
  }
  return de * kChargeFactor;
}

__device__ inline bool lanes_any(bool p) { return __ballot(p) != 0ull; }

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}

// Minimum-image shift (cubic rule of minimum_image_separation.f90:35-60) of
// one separation component.
__device__ inline double image_shift(double d, double hl, double l) { return d > hl ? -l : (d < -hl ? l : 0.0); }

// Pair terms of the placed molecule (px, py, pz) against frame molecule m:
// Compute_MoleculePair_Energy (energy_routines.f90:1589-1730) at the full
// cutoff, with the CBMC-cutoff sum taken in the same pass.  UNI: the image
// shift of every frame site is the one of the wavefront centre (c), valid
// when (wave extent + pair cutoff) < half the box: then every pair inside the
// cutoff gets exactly the per-pair minimum image, and every other pair stays
// outside it.  The pair code runs only when some lane of the wave has the
// pair inside rc_pair.
template <typename P, int NS, bool GEN, bool UNI>
__device__ inline void molecule_pairs(const FrameMol &m, const WidomFrame &f, const WidomPairTab &pt, const BoxDev &b,
                                      const WidomArgs &w, const PairConst<double> &cd, const PairConst<float> &cf,
                                      int coul_cbmc, double cx, double cy, double cz, bool live, const double (&px)[NS],
                                      const double (&py)[NS], const double (&pz)[NS], double &ev, double &eq,
                                      double &ecb, bool &over) {
  using TV = typename P::V;
  using TQ = typename P::Q;
  using M = typename P::M;
  const PairConst<TV> &cv = pconst<TV>(cd, cf);
  const PairConst<TQ> &cq = pconst<TQ>(cd, cf);
  for (int k = m.first; k < m.first + m.n; ++k) {
    const double4 st = f.site[k];
    const int tj = f.type[k];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (UNI) {
      sx = image_shift(cx - st.x, b.hL[0], b.L[0]);
      sy = image_shift(cy - st.y, b.hL[1], b.L[1]);
      sz = image_shift(cz - st.z, b.hL[2], b.L[2]);
    }
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      double dx = px[a] - st.x, dy = py[a] - st.y, dz = pz[a] - st.z;
      if (UNI) {
        dx += sx;
        dy += sy;
        dz += sz;
      } else {
        min_image(b, dx, dy, dz);
      }
      const double r2 = dx * dx + dy * dy + dz * dz;
      const bool near = live && (w.cutoff_only ? r2 <= w.rc_pair2 : true);
      if (!lanes_any(near)) continue;
      if (near) {
        if (r2 < cd.rc_low2) over = true;
        bool gv, gq;
        pair_cutoff(cd, false, r2, gv, gq);
        const bool incb = r2 <= cd.rc_cbmc2;
        const bool lj = tj >= 0 && w.type[a] >= 0 && cd.vdw_style != CASS_VDW_NONE;
        const double4 prm = lj ? pt.p[a * pt.ntypes + tj] : make_double4(0.0, 0.0, 0.0, 0.0);
        const bool do_v = lj && prm.x != 0.0;  // eps = 0 pairs add exactly 0
        const bool do_q = cd.q_style != CASS_Q_NONE && w.q[a] != 0.0 && st.w != 0.0;
        TV e_v = (TV)0;
        TQ e_q = (TQ)0;
        if (GEN) {
          if (do_v && (gv || incb)) e_v = vdw_inter<TV, M>(cv, (TV)r2, (TV)prm.x, (TV)prm.y, (TV)prm.z, (TV)prm.w);
          if (do_q && (gq || incb)) e_q = qq_inter<TQ, M>(cq, (TQ)r2, (TQ)(w.q[a] * st.w));
        } else {
          // LJ cut / cut_tail / cut_shift; Ewald real space or plain cut Coulomb
          if (do_v && (gv || incb)) e_v = lj_cut<TV, M>(cv, (TV)r2, (TV)prm.x, (TV)prm.y);
          if (do_q && (gq || incb)) {
            if (cd.q_style == CASS_Q_EWALD)
              e_q = ewald_real<TQ, M>(cq, (TQ)r2, (TQ)(w.q[a] * st.w));
            else
              e_q = coul_cut<TQ, M>((TQ)r2, (TQ)(w.q[a] * st.w));
          }
        }
        if (gv) ev += (double)e_v;
        if (gq) eq += (double)e_q;
        if (incb) ecb += (double)e_v + (coul_cbmc ? (double)e_q : 0.0);
      }
    }
  }
}

template <typename P, int NS, bool GEN>
__global__ void __launch_bounds__(kWBlock) k_widom_sorted(WidomArgs w, BoxDev b, WidomFrame f, WidomPairTab pt,
                                                          EwaldDev ew, PairConst<double> cd, PairConst<float> cf,
                                                          int coul_cbmc, const int *__restrict__ order,
                                                          double *__restrict__ part, int64_t *__restrict__ part_ov,
                                                          double *__restrict__ wvar, double *__restrict__ eout) {
  const int64_t s = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
  const bool active = s < w.n_ins;
  const int64_t i = order[active ? s : w.n_ins - 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double px[NS], py[NS], pz[NS];
  if (w.place) {
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      px[a] = w.place[(i * NS + a) * 3];
      py[a] = w.place[(i * NS + a) * 3 + 1];
      pz[a] = w.place[(i * NS + a) * 3 + 2];
    }
  } else {
    place_probe_n<NS>(w, b, (uint64_t)i, px, py, pz);
  }
  // Get_COM / Compute_Max_Com_Distance of the placed molecule
  double4 pc;
  {
    double tm = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      tm = tm + w.mass[a];
      cx = cx + w.mass[a] * px[a];
      cy = cy + w.mass[a] * py[a];
      cz = cz + w.mass[a] * pz[a];
    }
    cx = cx / tm;
    cy = cy / tm;
    cz = cz / tm;
    double dmax = 0.0;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      const double dx = px[a] - cx, dy = py[a] - cy, dz = pz[a] - cz;
      dmax = fmax(dmax, dx * dx + dy * dy + dz * dz);
    }
    pc = make_double4(cx, cy, cz, sqrt(dmax));
  }

  double ev = 0.0, eq = 0.0, ecb = 0.0;
  bool over = false;
  bool dead = !active;
  if (w.cutoff_only && b.cubic) {
    // Wave centre c (the first lane's centre of mass) and the extent of the
    // wave's sites around it.  Every pair contribution needs r <= rc_pair,
    // so a frame molecule is a candidate only if its centre lies within
    // rc_pair + max_dcom + rad of c; the whole wave skips the others.
    const double cx = __shfl(pc.x, 0, 64), cy = __shfl(pc.y, 0, 64), cz = __shfl(pc.z, 0, 64);
    double ext = 0.0, rad2 = 0.0;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      double dx = px[a] - cx, dy = py[a] - cy, dz = pz[a] - cz;
      // ext: raw coordinate differences (the shift argument needs them);
      // rad: periodic distance (the candidate test needs that)
      ext = fmax(ext, fmax(fabs(dx), fmax(fabs(dy), fabs(dz))));
      min_image(b, dx, dy, dz);
      rad2 = fmax(rad2, dx * dx + dy * dy + dz * dz);
    }
    ext = wave_max(ext);
    const double rad = sqrt(wave_max(rad2));
    const double hmin = fmin(b.hL[0], fmin(b.hL[1], b.hL[2]));
    const bool uni = ext + w.rc_pair < hmin;
    for (int base = 0; base < f.nmols && lanes_any(!dead); base += 64) {
      const int jt = base + lane;
      bool cand = false;
      if (jt < f.nmols) {
        const double4 cj = f.mol[jt].com;
        double dx = cj.x - cx, dy = cj.y - cy, dz = cj.z - cz;
        min_image(b, dx, dy, dz);
        const double reach = w.rc_pair + cj.w + rad + 1e-9;
        cand = dx * dx + dy * dy + dz * dz <= reach * reach;
      }
      uint64_t mask = __ballot(cand);
      while (mask) {
        const int j = base + __builtin_ctzll(mask);
        mask &= mask - 1;
        const FrameMol m = f.mol[j];
        if (uni)
          molecule_pairs<P, NS, GEN, true>(m, f, pt, b, w, cd, cf, coul_cbmc, cx, cy, cz, !dead, px, py, pz, ev, eq, ecb,
                                      over);
        else
          molecule_pairs<P, NS, GEN, false>(m, f, pt, b, w, cd, cf, coul_cbmc, cx, cy, cz, !dead, px, py, pz, ev, eq,
                                       ecb, over);
        dead = dead || over;
        if (!lanes_any(!dead)) break;
      }
    }
  } else {
    // every molecule, with the per-pair centre-of-mass test of
    // Check_MoleculePair_Cutoff (energy_routines.f90:3027-3088)
    for (int j = 0; j < f.nmols; ++j) {
      const FrameMol m = f.mol[j];
      const bool inr = !dead && (molpair_in_range(b, false, pc, m.com) || molpair_in_range(b, true, pc, m.com));
      if (!lanes_any(inr)) continue;
      molecule_pairs<P, NS, GEN, false>(m, f, pt, b, w, cd, cf, coul_cbmc, 0.0, 0.0, 0.0, inr, px, py, pz, ev, eq, ecb,
                                   over);
      dead = dead || over;
      if (!lanes_any(!dead)) break;
    }
  }
  const double e_inter = ev + eq;
  double de_recip = 0.0;
  if (w.charged && lanes_any(!dead)) {
    const double d = recip_delta_n<NS>(ew, b, w, px, py, pz);
    if (!over) de_recip = d;
  }
  const double dE = e_inter + de_recip + w.e_const;
  // Build_Molecule marks a trial whose CBMC-cutoff energy reaches max_kBT as
  // overlapped (fragment_growth.f90:536-539)
  const bool cbmc_over = over || (w.beta * ecb >= kMaxKBT);
  double wv = 0.0;
  int64_t ovl = 0;
  if (active) {
    if (!cbmc_over) wv = exp(-w.beta * dE);
    ovl = cbmc_over ? 1 : 0;
    if (wvar) wvar[i] = wv;
    if (eout) {
      eout[4 * i] = e_inter;
      eout[4 * i + 1] = ecb;
      eout[4 * i + 2] = de_recip;
      eout[4 * i + 3] = over ? 1.0 : 0.0;
    }
  }
  // block sums of widom_var and of the overlap count
  __shared__ double shv[kWBlock / 64];
  __shared__ int64_t sho[kWBlock / 64];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    wv += __shfl_down(wv, d, 64);
    ovl += __shfl_down(ovl, d, 64);
  }
  if (lane == 0) {
    shv[wid] = wv;
    sho[wid] = ovl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    int64_t c = 0;
#pragma unroll
    for (int k = 0; k < kWBlock / 64; ++k) {
      a += shv[k];
      c += sho[k];
    }
    part[blockIdx.x] = a;
    part_ov[blockIdx.x] = c;
  }
}

typedef void (*widom_kernel_t)(WidomArgs, BoxDev, WidomFrame, WidomPairTab, EwaldDev, PairConst<double>,
                               PairConst<float>, int, const int *, double *, int64_t *, double *, double *);

template <typename P, bool GEN>
static widom_kernel_t widom_kernel_for(int ns) {
  switch (ns) {
    case 1: return k_widom_sorted<P, 1, GEN>;
    case 2: return k_widom_sorted<P, 2, GEN>;
    case 3: return k_widom_sorted<P, 3, GEN>;
    case 4: return k_widom_sorted<P, 4, GEN>;
    case 5: return k_widom_sorted<P, 5, GEN>;
    case 6: return k_widom_sorted<P, 6, GEN>;
    case 7: return k_widom_sorted<P, 7, GEN>;
    case 8: return k_widom_sorted<P, 8, GEN>;
    default: return nullptr;
  }
}

// GEN: every vdW and Coulomb style; otherwise the common ones only (LJ cut,
// cut_tail, cut_shift; Ewald, cut Coulomb, none), whose pair code needs
// fewer registers.
static widom_kernel_t widom_kernel(int prec, int ns, bool gen) {
  if (prec == CASS_PREC_MIXED) return gen ? widom_kernel_for<PrecMixed, true>(ns) : widom_kernel_for<PrecMixed, false>(ns);
  if (prec == CASS_PREC_FP32) return gen ? widom_kernel_for<PrecFP32, true>(ns) : widom_kernel_for<PrecFP32, false>(ns);
  return gen ? widom_kernel_for<PrecFP64, true>(ns) : widom_kernel_for<PrecFP64, false>(ns);
}

}  // namespace cass

using namespace cass;

BoxDev cass_box_view(const cass_engine *e, const cass_box_state &b);  // engine.hip
EwaldDev cass_ewald_view(const cass_box_state &b);
int cass_widom_prepare(cass_engine *e, cass_box_state &b, int species, WidomArgs &w);  // host.cpp

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Cell grid of the insertion sort: about 64 insertions per cell, at most
// 64 cells per box edge.
static void widom_grid(const cass_box_state &b, int64_t n, int g[3]) {
  const double L[3] = {b.p.basis_length[0], b.p.basis_length[1], b.p.basis_length[2]};
  const double vol = std::max(L[0] * L[1] * L[2], 1e-30);
  const double cells = std::max(1.0, (double)n / 64.0);
  const double side = std::cbrt(vol / cells);
  for (int d = 0; d < 3; ++d) g[d] = (int)std::min(64.0, std::max(1.0, std::floor(L[d] / side + 0.5)));
}

int cass_run_widom(cass_engine *e, int32_t box, int32_t species, int64_t n_ins, uint64_t seed, const double *place,
                   double *widom_var, double *eout, cass_widom_result *res) {
  cass_box_state *bp = nullptr;
  if (!e || box < 0 || box >= (int)e->boxes.size()) return CASS_ERR_ARG;
  bp = &e->boxes[box];
  cass_box_state &b = *bp;
  if (species < 0 || species >= e->nspecies || n_ins <= 0 || n_ins > (int64_t)INT32_MAX) return CASS_ERR_ARG;
  const int ns = e->sp_natoms[species];
  if (ns > kMaxWidomSites || ns < 1) return cass_fail(e, CASS_ERR_ARG, "test molecule too large");
  WidomArgs w;
  int rc = cass_widom_prepare(e, b, species, w);
  if (rc) return rc;
  w.n_ins = n_ins;
  w.seed = seed;
  // every pair contribution of these styles needs r <= rc_pair
  // (Check_AtomPair_Cutoff, energy_routines.f90:3092-3195)
  {
    const cass_params &p = b.p;
    const bool vdw_cut = p.vdw_style == CASS_VDW_NONE ||
                         ((p.vdw_style == CASS_VDW_LJ || p.vdw_style == CASS_VDW_MIE) &&
                          (p.vdw_sum_style == CASS_SUM_CUT || p.vdw_sum_style == CASS_SUM_CUT_TAIL ||
                           p.vdw_sum_style == CASS_SUM_CUT_SHIFT || p.vdw_sum_style == CASS_SUM_CUT_SWITCH ||
                           p.vdw_sum_style == CASS_SUM_CUT_SHIFT_FORCE));
    const bool q_cut = p.charge_style == CASS_Q_NONE || p.charge_style == CASS_Q_EWALD ||
                       p.charge_style == CASS_Q_DSF || p.charge_style == CASS_Q_CUT;
    const double rc2 = std::max(std::max(p.rcut_vdwsq, p.rcut_coulsq),
                                std::max(p.rcut_cbmcsq, p.vdw_sum_style == CASS_SUM_CUT_SWITCH ? p.roff_switch_sq : 0.0));
    w.cutoff_only = (vdw_cut && q_cut) ? 1 : 0;
    w.rc_pair2 = rc2;
    w.rc_pair = std::sqrt(rc2);
  }

  // frame tables (host mirrors -> device)
  std::vector<FrameMol> fm;
  std::vector<double4> fs;
  std::vector<int> ft;
  fm.reserve(b.nmols);
  for (int m = 0; m < b.nmols; ++m) {
    if (!b.live[m]) continue;
    const int sp = b.mol_species[m], nat = e->sp_natoms[sp], off = e->sp_atom_off[sp];
    FrameMol r;
    r.com = b.comd_h[m];
    r.first = (int)fs.size();
    r.n = nat;
    r.pad0 = r.pad1 = 0;
    fm.push_back(r);
    const double *xyz = b.xyz_h.data() + 3 * (size_t)b.mol_first[m];
    for (int a = 0; a < nat; ++a) {
      fs.push_back(make_double4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], e->atom_charge[off + a]));
      ft.push_back(e->atom_type[off + a]);
    }
  }
  const int nt = e->ntypes;
  std::vector<double4> tab((size_t)ns * std::max(nt, 1));
  for (int a = 0; a < ns; ++a)
    for (int t = 0; t < nt; ++t) {
      const int ti = w.type[a];
      if (ti < 0) {
        tab[(size_t)a * nt + t] = make_double4(0.0, 0.0, 0.0, 0.0);
        continue;
      }
      const size_t k = (size_t)ti * nt + t;
      tab[(size_t)a * nt + t] = make_double4(e->eps[k], e->sig[k], e->mie_n[k], e->mie_m[k]);
    }

  const int nblk = (int)((n_ins + kWBlock - 1) / kWBlock);
  int g[3];
  widom_grid(b, n_ins, g);
  const int ncell = g[0] * g[1] * g[2];
  int bits = 1;
  while ((1 << bits) < ncell) ++bits;
  size_t sort_tmp = 0;
  HIP_TRY(e, rocprim::radix_sort_pairs(nullptr, sort_tmp, (unsigned *)nullptr, (unsigned *)nullptr, (int *)nullptr,
                                       (int *)nullptr, (unsigned)n_ins, 0, bits, e->stream));
  const size_t n = (size_t)n_ins;
  size_t off = 0;
  const size_t o_part = off; off = align256(off + sizeof(double) * nblk);
  const size_t o_pov = off; off = align256(off + sizeof(int64_t) * nblk);
  const size_t o_sum = off; off = align256(off + 64);
  const size_t o_k0 = off; off = align256(off + sizeof(unsigned) * n);
  const size_t o_k1 = off; off = align256(off + sizeof(unsigned) * n);
  const size_t o_v0 = off; off = align256(off + sizeof(int) * n);
  const size_t o_v1 = off; off = align256(off + sizeof(int) * n);
  const size_t o_fm = off; off = align256(off + sizeof(FrameMol) * std::max<size_t>(fm.size(), 1));
  const size_t o_fs = off; off = align256(off + sizeof(double4) * std::max<size_t>(fs.size(), 1));
  const size_t o_ft = off; off = align256(off + sizeof(int) * std::max<size_t>(ft.size(), 1));
  const size_t o_tab = off; off = align256(off + sizeof(double4) * tab.size());
  const size_t o_tmp = off; off = align256(off + sort_tmp);
  char *s = (char *)cass_scratch(e, off);
  if (!s) return cass_fail(e, CASS_ERR_NOMEM, "scratch");
  double *part = (double *)(s + o_part);
  int64_t *part_ov = (int64_t *)(s + o_pov);
  double *d_sum = (double *)(s + o_sum);
  int64_t *d_nov = (int64_t *)(d_sum + 1);
  unsigned *k0 = (unsigned *)(s + o_k0), *k1 = (unsigned *)(s + o_k1);
  int *v0 = (int *)(s + o_v0), *v1 = (int *)(s + o_v1);
  FrameMol *d_fm = (FrameMol *)(s + o_fm);
  double4 *d_fs = (double4 *)(s + o_fs);
  int *d_ft = (int *)(s + o_ft);
  double4 *d_tab = (double4 *)(s + o_tab);
  if (!fm.empty()) {
    HIP_TRY(e, hipMemcpyAsync(d_fm, fm.data(), sizeof(FrameMol) * fm.size(), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(d_fs, fs.data(), sizeof(double4) * fs.size(), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(d_ft, ft.data(), sizeof(int) * ft.size(), hipMemcpyHostToDevice, e->stream));
  }
  HIP_TRY(e, hipMemcpyAsync(d_tab, tab.data(), sizeof(double4) * tab.size(), hipMemcpyHostToDevice, e->stream));

  double *d_place = nullptr, *d_var = nullptr, *d_eout = nullptr;
  if (place) {
    HIP_TRY(e, hipMallocAsync((void **)&d_place, sizeof(double) * n * ns * 3, e->stream));
    HIP_TRY(e, hipMemcpyAsync(d_place, place, sizeof(double) * n * ns * 3, hipMemcpyHostToDevice, e->stream));
  }
  if (widom_var) HIP_TRY(e, hipMallocAsync((void **)&d_var, sizeof(double) * n, e->stream));
  if (eout) HIP_TRY(e, hipMallocAsync((void **)&d_eout, sizeof(double) * 4 * n, e->stream));
  w.place = d_place;

  BoxDev bv = cass_box_view(e, b);
  EwaldDev ew = cass_ewald_view(b);
  if (!b.ewald_ready) {
    ew.nvecs = 0;
    ew.nrows = 0;
  }
  PairConst<double> cd = make_pair_const<double>(b.p, e->ntypes);
  PairConst<float> cf = make_pair_const<float>(b.p, e->ntypes);
  WidomFrame wf{d_fm, d_fs, d_ft, (int)fm.size()};
  WidomPairTab pt{d_tab, nt};
  const int coul_cbmc = e->sp_coul_cbmc[species];

  // spatial order of the insertions
  hipLaunchKernelGGL(k_widom_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, w, bv, g[0], g[1],
                     g[2], k0, v0);
  HIP_TRY(e, hipGetLastError());
  HIP_TRY(e, rocprim::radix_sort_pairs((void *)(s + o_tmp), sort_tmp, k0, k1, v0, v1, (unsigned)n, 0, bits,
                                       e->stream));
  const cass_params &pp = b.p;
  const bool common = (pp.vdw_style == CASS_VDW_NONE ||
                       (pp.vdw_style == CASS_VDW_LJ &&
                        (pp.vdw_sum_style == CASS_SUM_CUT || pp.vdw_sum_style == CASS_SUM_CUT_TAIL ||
                         pp.vdw_sum_style == CASS_SUM_CUT_SHIFT))) &&
                      (pp.charge_style == CASS_Q_NONE || pp.charge_style == CASS_Q_EWALD ||
                       pp.charge_style == CASS_Q_CUT);
  widom_kernel_t kern = widom_kernel(e->prec, ns, !common);
  HIP_TRY(e, hipEventRecord(e->ev0, e->stream));
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(kWBlock), 0, e->stream, w, bv, wf, pt, ew, cd, cf, coul_cbmc, v1, part,
                     part_ov, d_var, d_eout);
  HIP_TRY(e, hipGetLastError());
  HIP_TRY(e, hipEventRecord(e->ev1, e->stream));
  hipLaunchKernelGGL(k_sum_widom, dim3(1), dim3(256), 0, e->stream, part, part_ov, nblk, d_sum, d_nov);
  HIP_TRY(e, hipGetLastError());
  double hs[2];
  HIP_TRY(e, hipMemcpyAsync(hs, d_sum, sizeof(double) * 2, hipMemcpyDeviceToHost, e->stream));
  if (widom_var) HIP_TRY(e, hipMemcpyAsync(widom_var, d_var, sizeof(double) * n, hipMemcpyDeviceToHost, e->stream));
  if (eout) HIP_TRY(e, hipMemcpyAsync(eout, d_eout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, e->stream));
  if (d_place) HIP_TRY(e, hipFreeAsync(d_place, e->stream));
  if (d_var) HIP_TRY(e, hipFreeAsync(d_var, e->stream));
  if (d_eout) HIP_TRY(e, hipFreeAsync(d_eout, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  float ms = 0.0f;
  HIP_TRY(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->last_kernel_ms = ms;
  res->widom_sum = hs[0];
  std::memcpy(&res->n_overlaps, &hs[1], sizeof(int64_t));
  res->n_insertions = n_ins;
  res->e_const = w.e_const;
  return CASS_OK;
}
