// nvtdev.hip -- the NVT move loop of cass_nvt_run resident on the device.
//
// nvtmc_driver.f90:94-152 choosing between Translate (move_translate.f90:23-360)
// and Rotate (move_rotate.f90:23-520), with the same random-number draws
// (random_generators.f90:51-122), energy bookkeeping and accept_or_reject
// (accept_or_reject.f90:22-65) as the host-driven loop in nvt.cpp.  One
// workgroup of 1024 threads runs the moves in sequence: thread 0 draws the
// move and builds the trial coordinates, all threads share the molecule-pair
// energies (old and new placements side by side) and the k vectors of the
// reciprocal update, and an accepted move is committed in device memory.  No
// host round trip per move: the host launches chunks of moves and reads the
// RNG state, energies and records back at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace cass {

constexpr int kNvtBlock = 1024;
constexpr int kNvtWaves = kNvtBlock / 64;

struct NvtChainArgs {
  BoxDev b;
  TopoDev tp;
  PairConst<double> cd;
  PairConst<float> cf;
  EwaldDev ew;
  double *x, *y, *z;  // the box's site coordinates (the memory of b.x, b.y, b.z)
  double4 *comd;      // the box's (COM, max_dcom) (the memory of b.comd)
  const int *mols;    // mobile molecules in locate order
  int nm, species, n, spa0, charged, linear, coul_cbmc;
  double prob_translate, max_disp, max_rot, beta;
  int64_t nsteps, step0;  // moves of this launch; index of its first move
  int64_t *rng;           // 5 words, in/out
  double *en;             // total, inter, inter_vdw, inter_q, reciprocal; in/out
  uint8_t *kind, *accepted;  // records per move (index step0 + s), or null
  double *energy, *ln_pacc;
  int64_t *rng_state;
  int *status;  // set to 1 when a molecule overlaps before its move
};

__device__ inline int64_t nvt_lsl(int64_t v, int k) { return (int64_t)((uint64_t)v << k); }
__device__ inline int64_t nvt_lsr(int64_t v, int k) { return (int64_t)((uint64_t)v >> k); }

// rranf (random_generators.f90:51-122): five-component LFSR, u in (0, 1)
__device__ inline double nvt_uniform(int64_t *s) {
  int64_t b;
  b = nvt_lsr(nvt_lsl(s[0], 1) ^ s[0], 53);
  s[0] = nvt_lsl(s[0] & -2, 10) ^ b;
  b = nvt_lsr(nvt_lsl(s[1], 24) ^ s[1], 50);
  s[1] = nvt_lsl(s[1] & -512, 5) ^ b;
  b = nvt_lsr(nvt_lsl(s[2], 3) ^ s[2], 23);
  s[2] = nvt_lsl(s[2] & -4096, 29) ^ b;
  b = nvt_lsr(nvt_lsl(s[3], 5) ^ s[3], 24);
  s[3] = nvt_lsl(s[3] & -131072, 23) ^ b;
  b = nvt_lsr(nvt_lsl(s[4], 3) ^ s[4], 33);
  s[4] = nvt_lsl(s[4] & -8388608, 8) ^ b;
  return (double)(s[0] ^ s[1] ^ s[2] ^ s[3] ^ s[4]) * 5.4210108624275221E-20 + 0.5;
}

__device__ inline double nvt_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ inline int nvt_wave_or(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_down(v, o, 64);
  return v;
}

// Workgroup sum of NV doubles and an OR of flags in a fixed order (waves,
// then the waves in index order); every thread receives the result.
template <int NV>
__device__ inline void nvt_reduce(double (&v)[NV], int &flag, double (*sh)[kNvtWaves], int *shf, double *res,
                                  int *resf) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = nvt_wave_sum(v[k]);
  flag = nvt_wave_or(flag);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
    shf[wid] = flag;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kNvtWaves; ++w) s += sh[k][w];
      res[k] = s;
    }
    int f = 0;
    for (int w = 0; w < kNvtWaves; ++w) f |= shf[w];
    *resf = f;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = res[k];
  flag = *resf;
}

// One axis of Fold_Molecule for a cubic box: the centre c and the sites q
// move by one box length when c leaves [-h, h].
__device__ inline void nvt_fold(double &c, double *q, int n, double L, double h) {
  if (c > h) {
    c = c - L;
    for (int i = 0; i < n; ++i) q[i] = q[i] - L;
  } else if (c < -h) {
    c = c + L;
    for (int i = 0; i < n; ++i) q[i] = q[i] + L;
  }
}

__global__ void __launch_bounds__(kNvtBlock) k_nvt_chain(NvtChainArgs a) {
  __shared__ double ox[kMaxSites], oy[kMaxSites], oz[kMaxSites];
  __shared__ double qx[kMaxSites], qy[kMaxSites], qz[kMaxSites];
  __shared__ int pspa[kMaxSites];
  __shared__ double4 com_old, com_new, cn_fold;
  __shared__ int s_m, s_rot, s_acc;
  __shared__ double sh[4][kNvtWaves];
  __shared__ int shf[kNvtWaves];
  __shared__ double res[4];
  __shared__ int resf;
  const int tid = threadIdx.x;
  const int n = a.n, K = a.ew.nvecs, nmols = a.b.nmols;
  int64_t rs[5];
  double en[5];
  if (tid == 0) {
    for (int k = 0; k < 5; ++k) rs[k] = a.rng[k];
    for (int k = 0; k < 5; ++k) en[k] = a.en[k];
  }
  if (tid < n) pspa[tid] = a.spa0 + tid;
  __syncthreads();
  for (int64_t st = 0; st < a.nsteps; ++st) {
    // --- the move (thread 0): draws in the order of nvtmc_driver / Translate / Rotate
    if (tid == 0) {
      const double rno = nvt_uniform(rs);
      const int rot = !(rno <= a.prob_translate);
      (void)nvt_uniform(rs);  // species pick (one mobile species)
      const int im = (int)(nvt_uniform(rs) * a.nm);
      const int m = a.mols[im];
      const int f = a.b.first[m];
      const double4 co = a.comd[m];
      double4 cn = co;
      for (int i = 0; i < n; ++i) {
        ox[i] = a.x[f + i];
        oy[i] = a.y[f + i];
        oz[i] = a.z[f + i];
      }
      if (!rot) {
        const double dx = (2.0 * nvt_uniform(rs) - 1.0) * a.max_disp;
        const double dy = (2.0 * nvt_uniform(rs) - 1.0) * a.max_disp;
        const double dz = (2.0 * nvt_uniform(rs) - 1.0) * a.max_disp;
        for (int i = 0; i < n; ++i) {
          qx[i] = ox[i] + dx;
          qy[i] = oy[i] + dy;
          qz[i] = oz[i] + dz;
        }
        cn.x = co.x + dx;
        cn.y = co.y + dy;
        cn.z = co.z + dz;
        // Fold_Molecule (minimum_image_separation.f90:162-240), cubic boxes
        if (a.b.cubic) {
          nvt_fold(cn.x, qx, n, a.b.L[0], a.b.hL[0]);
          nvt_fold(cn.y, qy, n, a.b.L[1], a.b.hL[1]);
          nvt_fold(cn.z, qz, n, a.b.L[2], a.b.hL[2]);
        }
      } else {
        double r[9];
        if (a.linear) {
          // Rotate_Molecule_Eulerian (move_rotate.f90:468-520)
          const double th = acos(1.0 - 2.0 * nvt_uniform(rs));
          const double ph = (1.0 - 2.0 * nvt_uniform(rs)) * kPi;
          const double ps = (1.0 - 2.0 * nvt_uniform(rs)) * kPi;
          r[0] = cos(ps) * cos(ph) - cos(th) * sin(ph) * sin(ps);
          r[1] = cos(ps) * sin(ph) + cos(th) * cos(ph) * sin(ps);
          r[2] = sin(ps) * sin(th);
          r[3] = -sin(ps) * cos(ph) - cos(th) * sin(ph) * cos(ps);
          r[4] = -sin(ps) * sin(ph) + cos(th) * cos(ph) * cos(ps);
          r[5] = cos(ps) * sin(th);
          r[6] = sin(th) * sin(ph);
          r[7] = -sin(th) * cos(ph);
          r[8] = cos(th);
        } else {
          // Rotate_Molecule_Axis (move_rotate.f90:362-463)
          double p1 = 0.0, p2 = 0.0, p3 = 0.0;
          const int axis = (int)(3.0 * nvt_uniform(rs)) + 1;
          const double ang = (2.0 * nvt_uniform(rs) - 1.0) * a.max_rot;
          if (axis == 1)
            p1 = ang;
          else if (axis == 2)
            p2 = ang;
          else
            p3 = ang;
          const double c1 = cos(p1), c2 = cos(p2), c3 = cos(p3);
          const double s1 = sin(p1), s2 = sin(p2), s3 = sin(p3);
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
        for (int i = 0; i < n; ++i) {
          const double px = ox[i] - co.x, py = oy[i] - co.y, pz = oz[i] - co.z;
          qx[i] = (r[0] * px + r[1] * py + r[2] * pz) + co.x;
          qy[i] = (r[3] * px + r[4] * py + r[5] * pz) + co.y;
          qz[i] = (r[6] * px + r[7] * py + r[8] * pz) + co.z;
        }
      }
      s_m = m;
      s_rot = rot;
      cn_fold = cn;
    }
    __syncthreads();
    // the trial kernel's prefilter centres: Get_COM of each placement
    if (tid == 0) com_old = molecule_com(a.tp, a.spa0, n, ox, oy, oz);
    if (tid == 64) com_new = molecule_com(a.tp, a.spa0, n, qx, qy, qz);
    __syncthreads();
    const int m = s_m;
    // --- inter energies of the old and the new placement (Compute_Molecule_Nonbond_Inter_Energy)
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    int ov = 0;
    for (int t = tid; t < 2 * nmols; t += kNvtBlock) {
      const int which = t >= nmols ? 1 : 0;
      const int j = t - which * nmols;
      if (j == m || !a.b.live[j]) continue;
      if (!molpair_in_range(a.b, false, which ? com_new : com_old, a.b.comd[j])) continue;
      double ev = 0.0, eq = 0.0;
      const bool o = molpair_energy<PrecFP64>(a.b, a.tp, a.cd, a.cf, false, a.coul_cbmc != 0, n, which ? qx : ox,
                                              which ? qy : oy, which ? qz : oz, pspa, j, ev, eq, nullptr);
      if (o) {
        ov |= 1 << which;
      } else {
        v[2 * which] += ev;
        v[2 * which + 1] += eq;
      }
    }
    nvt_reduce<4>(v, ov, sh, shf, res, &resf);
    if (ov & 1) {
      if (tid == 0) *a.status = 1;
      __syncthreads();
      break;
    }
    // --- reciprocal space: Update_System_Ewald_Reciprocal_Energy (staged)
    double erm = 0.0;
    const bool reach = !(ov & 2);
    if (reach && a.charged) {
      double e1[1] = {0.0};
      int f0 = 0;
      for (int k = tid; k < K; k += kNvtBlock) {
        const size_t mk = (size_t)m * K + k;
        const double hx = a.ew.hx[k], hy = a.ew.hy[k], hz = a.ew.hz[k];
        double c = 0.0, s = 0.0;
        for (int i = 0; i < n; ++i) {
          const double hr = hx * qx[i] + hy * qy[i] + hz * qz[i];
          const double q = a.tp.atom_charge[a.spa0 + i];
          double sv, cv;
          sincos(hr, &sv, &cv);
          c = c + q * cv;
          s = s + q * sv;
        }
        const double c1 = a.ew.cs[k] + (c - a.ew.cmol[mk]), s1 = a.ew.sn[k] + (s - a.ew.smol[mk]);
        a.ew.cs_new[k] = c1;
        a.ew.sn_new[k] = s1;
        a.ew.cm_new[k] = c;
        a.ew.sm_new[k] = s;
        e1[0] += a.ew.cn[k] * (c1 * c1 + s1 * s1);
      }
      nvt_reduce<1>(e1, f0, sh, shf, res, &resf);
      erm = e1[0] * kChargeFactor;
    }
    // --- acceptance (thread 0)
    if (tid == 0) {
      double lnp = std::numeric_limits<double>::quiet_NaN();
      int acc = 0;
      if (reach) {
        double dE = 0.0;
        if (a.charged) dE = erm - en[4];
        if (!s_rot)
          dE = dE + (v[2] - v[0]) + (v[3] - v[1]);
        else
          dE = dE + v[2] - v[0] + v[3] - v[1];
        lnp = a.beta * dE;
        if (lnp <= 0.0) {
          acc = 1;
        } else if (lnp < kMaxKBT) {
          const double pacc = exp(-lnp);
          if (nvt_uniform(rs) <= pacc) acc = 1;
        }
        if (acc) {
          en[0] = en[0] + dE;
          en[1] = en[1] + dE;
          en[2] = en[2] + v[2] - v[0];
          en[3] = en[3] + v[3] - v[1];
          if (a.charged) en[4] = erm;
          // the stored centre: the folded one after a translation, Get_COM
          // after a rotation (move_rotate.f90:287); max_dcom unchanged
          double4 c = a.comd[m];
          if (!s_rot) {
            c.x = cn_fold.x;
            c.y = cn_fold.y;
            c.z = cn_fold.z;
          } else {
            double tm = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
            for (int i = 0; i < n; ++i) {
              const double ms = a.tp.atom_mass[a.spa0 + i];
              tm = tm + ms;
              cx = cx + ms * qx[i];
              cy = cy + ms * qy[i];
              cz = cz + ms * qz[i];
            }
            c.x = cx / tm;
            c.y = cy / tm;
            c.z = cz / tm;
          }
          a.comd[m] = c;
        }
      }
      s_acc = acc;
      const int64_t g = a.step0 + st;
      if (a.kind) a.kind[g] = (uint8_t)s_rot;
      if (a.accepted) a.accepted[g] = (uint8_t)acc;
      if (a.energy) a.energy[g] = en[0];
      if (a.ln_pacc) a.ln_pacc[g] = lnp;
      if (a.rng_state)
        for (int k = 0; k < 5; ++k) a.rng_state[5 * g + k] = rs[k];
    }
    __syncthreads();
    if (s_acc) {
      const int f = a.b.first[m];
      if (tid < n) {
        a.x[f + tid] = qx[tid];
        a.y[f + tid] = qy[tid];
        a.z[f + tid] = qz[tid];
      }
      if (a.charged)
        for (int k = tid; k < K; k += kNvtBlock) {
          const size_t mk = (size_t)m * K + k;
          a.ew.cs[k] = a.ew.cs_new[k];
          a.ew.sn[k] = a.ew.sn_new[k];
          a.ew.cmol[mk] = a.ew.cm_new[k];
          a.ew.smol[mk] = a.ew.sm_new[k];
        }
    }
    __threadfence_block();
    __syncthreads();
  }
  if (tid == 0) {
    for (int k = 0; k < 5; ++k) a.rng[k] = rs[k];
    for (int k = 0; k < 5; ++k) a.en[k] = en[k];
  }
}

}  // namespace cass

using namespace cass;

BoxDev cass_box_view(const cass_engine *e, const cass_box_state &b);  // engine.hip
TopoDev cass_topo_view(const cass_engine *e);
EwaldDev cass_ewald_view(const cass_box_state &b);

extern "C" int cass_nvt_run_device(cass_engine *e, int32_t box, cass_rng *rng, const cass_nvt_options *opt,
                                   int64_t nsteps, cass_energy *en, cass_nvt_record *rec) {
  if (!e || !rng || !opt || !en || nsteps < 0 || box < 0 || box >= (int)e->boxes.size()) return CASS_ERR_ARG;
  cass_box_state &b = e->boxes[box];
  const int sp = opt->species;
  if (sp < 0 || sp >= e->nspecies) return CASS_ERR_ARG;
  const int n = e->sp_natoms[sp];
  if (n > kMaxSites) return cass_fail(e, CASS_ERR_ARG, "molecule has more sites than the device chain holds");
  bool has_q = false;
  for (int a = 0; a < n; ++a) has_q = has_q || e->atom_charge[e->sp_atom_off[sp] + a] != 0.0;
  const bool charged = b.p.charge_style == CASS_Q_EWALD && has_q;
  if (charged && !b.ewald_ready) return cass_fail(e, CASS_ERR_STATE, "Ewald sums not set up");
  std::vector<int> mols;  // locate order of the mobile species
  for (int m = 0; m < b.nmols; ++m)
    if (b.mol_species[m] == sp && b.live[m]) mols.push_back(m);
  const int nm = (int)mols.size();
  if (nm == 0) return cass_fail(e, CASS_ERR_STATE, "no molecules to move");
  if (nsteps == 0) return CASS_OK;
  HIP_TRY(e, hipSetDevice(e->device));
  HIP_TRY(e, hipStreamSynchronize(e->stream));

  // device state of the chain: RNG words, energies, status, mobile list, records
  const size_t ns = (size_t)nsteps;
  int64_t *d_rng = nullptr;
  HIP_TRY(e, hipMallocAsync((void **)&d_rng, sizeof(int64_t) * 5 + sizeof(double) * 5 + sizeof(int) * (2 + nm),
                            e->stream));
  double *d_en = (double *)(d_rng + 5);
  int *d_status = (int *)(d_en + 5);
  int *d_mols = d_status + 2;
  uint8_t *d_kind = nullptr, *d_acc = nullptr;
  double *d_energy = nullptr, *d_lnp = nullptr;
  int64_t *d_rs = nullptr;
  if (rec && rec->kind) HIP_TRY(e, hipMallocAsync((void **)&d_kind, ns, e->stream));
  if (rec && rec->accepted) HIP_TRY(e, hipMallocAsync((void **)&d_acc, ns, e->stream));
  if (rec && rec->energy) HIP_TRY(e, hipMallocAsync((void **)&d_energy, sizeof(double) * ns, e->stream));
  if (rec && rec->ln_pacc) HIP_TRY(e, hipMallocAsync((void **)&d_lnp, sizeof(double) * ns, e->stream));
  if (rec && rec->rng_state) HIP_TRY(e, hipMallocAsync((void **)&d_rs, sizeof(int64_t) * 5 * ns, e->stream));
  const double en_in[5] = {en->total, en->inter, en->inter_vdw, en->inter_q, en->reciprocal};
  const int zero[2] = {0, 0};
  HIP_TRY(e, hipMemcpyAsync(d_rng, rng->s, sizeof(int64_t) * 5, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipMemcpyAsync(d_en, en_in, sizeof(double) * 5, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipMemcpyAsync(d_status, zero, sizeof(int) * 2, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipMemcpyAsync(d_mols, mols.data(), sizeof(int) * nm, hipMemcpyHostToDevice, e->stream));

  NvtChainArgs a;
  std::memset(&a, 0, sizeof(a));
  a.b = cass_box_view(e, b);
  a.tp = cass_topo_view(e);
  a.cd = make_pair_const<double>(b.p, e->ntypes);
  a.cf = make_pair_const<float>(b.p, e->ntypes);
  a.ew = cass_ewald_view(b);
  if (!charged) a.ew.nvecs = 0;
  a.x = b.x;
  a.y = b.y;
  a.z = b.z;
  a.comd = b.comd;
  a.mols = d_mols;
  a.nm = nm;
  a.species = sp;
  a.n = n;
  a.spa0 = e->sp_atom_off[sp];
  a.charged = charged ? 1 : 0;
  a.linear = opt->linear;
  a.coul_cbmc = e->sp_coul_cbmc[sp];
  a.prob_translate = opt->prob_translate;
  a.max_disp = opt->max_disp;
  a.max_rot = opt->max_rot;
  a.beta = b.p.beta;
  a.rng = d_rng;
  a.en = d_en;
  a.kind = d_kind;
  a.accepted = d_acc;
  a.energy = d_energy;
  a.ln_pacc = d_lnp;
  a.rng_state = d_rs;
  a.status = d_status;
  // chunks of moves per launch (a launch runs a few tens of milliseconds)
  const int64_t chunk = 4096;
  for (int64_t s0 = 0; s0 < nsteps; s0 += chunk) {
    a.step0 = s0;
    a.nsteps = std::min(chunk, nsteps - s0);
    hipLaunchKernelGGL(k_nvt_chain, dim3(1), dim3(kNvtBlock), 0, e->stream, a);
    HIP_TRY(e, hipGetLastError());
  }
  int64_t rs_out[5];
  double en_out[5];
  int status[2];
  HIP_TRY(e, hipMemcpyAsync(rs_out, d_rng, sizeof(int64_t) * 5, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipMemcpyAsync(en_out, d_en, sizeof(double) * 5, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipMemcpyAsync(status, d_status, sizeof(int) * 2, hipMemcpyDeviceToHost, e->stream));
  if (d_kind) HIP_TRY(e, hipMemcpyAsync(rec->kind, d_kind, ns, hipMemcpyDeviceToHost, e->stream));
  if (d_acc) HIP_TRY(e, hipMemcpyAsync(rec->accepted, d_acc, ns, hipMemcpyDeviceToHost, e->stream));
  if (d_energy) HIP_TRY(e, hipMemcpyAsync(rec->energy, d_energy, sizeof(double) * ns, hipMemcpyDeviceToHost, e->stream));
  if (d_lnp) HIP_TRY(e, hipMemcpyAsync(rec->ln_pacc, d_lnp, sizeof(double) * ns, hipMemcpyDeviceToHost, e->stream));
  if (d_rs) HIP_TRY(e, hipMemcpyAsync(rec->rng_state, d_rs, sizeof(int64_t) * 5 * ns, hipMemcpyDeviceToHost, e->stream));
  // host mirrors of the moved molecules
  const size_t nsite = b.xyz_h.size() / 3;
  std::vector<double> hx(nsite), hy(nsite), hz(nsite);
  if (nsite) {
    HIP_TRY(e, hipMemcpyAsync(hx.data(), b.x, sizeof(double) * nsite, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipMemcpyAsync(hy.data(), b.y, sizeof(double) * nsite, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipMemcpyAsync(hz.data(), b.z, sizeof(double) * nsite, hipMemcpyDeviceToHost, e->stream));
  }
  if (b.nmols)
    HIP_TRY(e, hipMemcpyAsync(b.comd_h.data(), b.comd, sizeof(double4) * b.nmols, hipMemcpyDeviceToHost, e->stream));
  for (void *p : {(void *)d_kind, (void *)d_acc, (void *)d_energy, (void *)d_lnp, (void *)d_rs})
    if (p) HIP_TRY(e, hipFreeAsync(p, e->stream));
  HIP_TRY(e, hipFreeAsync(d_rng, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < nsite; ++i) {
    b.xyz_h[3 * i] = hx[i];
    b.xyz_h[3 * i + 1] = hy[i];
    b.xyz_h[3 * i + 2] = hz[i];
  }
  ++e->gen;
  std::memcpy(rng->s, rs_out, sizeof(rs_out));
  en->total = en_out[0];
  en->inter = en_out[1];
  en->inter_vdw = en_out[2];
  en->inter_q = en_out[3];
  en->reciprocal = en_out[4];
  if (charged) b.e_recip = en_out[4];
  b.staged_mol = -1;
  if (status[0]) return cass_fail(e, CASS_ERR_STATE, "molecule energy too high before the move");
  return CASS_OK;
}
