// adjoint.hip -- gradient of a far-field spectral loss with respect to every
// track sample and particle weight (gfx950).
//
// Forward (kernel_farfield.cl:59-106, one snapshot record r per particle):
//   S_g = sum_p w_p dt^2 |F_p(g)|^2,  F = sum_{s <= r, active} A_s(n) e^{i phi_s},
//   phi_s = omega (t_s - n.x_s),  A = c1 (n - m) - c2 a,  c2 = 1/(1 - n.m),
//   c1 = (n.a) c2^2,  m = (b_s + b_{s+1})/2,  a = (b_{s+1} - b_s)/dt,  b = u/sqrt(1+u.u).
// Given G = dL/dS (one array for 'total', three for 'cartesian') this adds
//   dL/dw_p = sum_g dt^2 G_g |F_p(g)|^2
//   dL/dx_s = n (A . R),   R_c = sum_omega omega H_c Im(conj(F_c) e^{i phi_s})
//   dL/dA_s = P,           P_c = sum_omega H_c Re(conj(F_c) e^{i phi_s}),  H_c = 2 w dt^2 G_c
//   dL/dm = 2 c1 c2 n ((n-m).P) - c1 P - c2^2 n (a.P),   dL/da = c2^2 n ((n-m).P) - c2 P
//   dL/db_s += dL/dm/2 - dL/da/dt,  dL/db_{s+1} += dL/dm/2 + dL/da/dt
//   dL/du = (dL/db - b (b . dL/db)) / gamma
// with the |dphase| < pi mask held fixed (it is piecewise constant).
//
// Layout.  The fields F_p(g) of a chunk of particles come from the spectrum
// tile kernel itself (an internal per-particle output, spectrum.hip), into a
// workspace of at most ~2 GiB; then a block owns (particle p, direction group
// q) and walks the group's directions.  Per direction it turns F into per-omega
// coefficients H_c F_c in LDS, then sweeps the transpose (threads = steps, loop
// over omegas) and adds the step's dL/dx, dL/dm, dL/da into its group's slab.
// Slabs are summed over groups in a fixed order by a second kernel, which also
// applies the m, a -> b -> u chain: no atomics, and the result does not depend
// on scheduling.
//
// Split precision, ascending non-negative linear grid: a step keeps exactly the
// omegas below its threshold |omega|/2pi < 1/(2 |tau_s - tau_{s-1}|) (the
// forward's test), a prefix of the table found by binary search; exp(i omega
// tau) walks the prefix by rotation with exp(i domega tau), restarted from an
// FP64 phase every 32 omegas.  'total' keeps the coefficients on e_theta,
// e_phi (F is perpendicular to n).  FP64 keeps the reference's exact per-pair
// test and an FP64 sin/cos per pair.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>
#include <math.h>

#include <algorithm>

#include "synchrad_amd.h"
#include "common.h"

namespace {

using namespace sr;

// The per-omega coefficient tables are held in LDS for at most this many
// omegas at a time; larger grids are swept in omega tiles (the step's P and R
// sums are sums over omega, and its gradient words are linear in them, so
// each tile adds its part)
constexpr int kAdjOmegaTile = 512;
constexpr int kSlabWords = 9;      // dL/dx (3), dL/dm (3), dL/da (3) per interval

struct AParams {
  // tracks (same field names as the spectrum kernel's, for load_particle)
  const int64_t *offsets;
  const double *x, *y, *z, *ux, *uy, *uz;
  const double *weight;
  const uint32_t *it_start, *it_end, *it_snaps;
  uint32_t n_snaps, snaps_stride;
  double dt;
  int64_t single_n;
  double single_w;
  uint32_t single_its, single_ite;
  int64_t n_particles, n_points;
  // grid
  int32_t comp;
  uint32_t n_omega, n_theta, n_phi, n_dir;
  const double *omega, *sin_t, *cos_t, *sin_p, *cos_p;
  double domega_rev;      // constant omega step / 2 pi, or 0
  const double *dS[3];
  // fields of particles [chunk_base, chunk_base + chunk_n): F[6][chunk_n][n_grid]
  const double *F;
  // step data, one record per sample [n_points][a_x, a_y, a_z, bbar_x, bbar_y, bbar_z]
  // (spectrum.hip step_prep)
  const double *pre;
  int64_t chunk_base, chunk_n, n_grid;
  // decomposition and workspace
  int32_t n_groups;
  int32_t dir_batch;   // directions whose coefficient tables share the block's LDS
  double *slab;    // [n_groups][kSlabWords][n_points]
  double *slab_w;  // [n_groups][n_particles]
  // outputs (+=)
  double *gx, *gy, *gz, *gux, *guy, *guz, *gw;
};

// frequencies per iteration of the transpose sweep's direct omega loop
#ifndef SR_ADJ_UNROLL
#define SR_ADJ_UNROLL 4
#endif
// and of the split sweep's rotation loop (the recurrence is serial; unrolling
// lets the next frequencies' table reads and sums overlap it)
#ifndef SR_ADJ_ROT_UNROLL
#define SR_ADJ_ROT_UNROLL 4
#endif

// TWO: 'total', coefficients on e_theta (word 0) and e_phi (word 1); the
// fields are transverse, so H F has no part along n
template <int PREC, bool TWO>
__global__ void __launch_bounds__(kMaxT) k_adjoint(const AParams p) {
  constexpr bool kSplit = (PREC == SR_PREC_SPLIT);
  using Acc = typename std::conditional<kSplit, float, double>::type;
  constexpr int NCW = TWO ? 2 : 3;   // coefficient words per omega: re 0..NCW-1, im NCW..2NCW-1
  constexpr int NW = 2 * NCW;
  __shared__ double s_cw[kAdjOmegaTile];     // the tile's omega (F64) or omega/2pi (split)
  __shared__ float s_cwf[kSplit ? kAdjOmegaTile : 1];   // split: s_cw as the sums' FP32 factor
  __shared__ double s_red[kMaxT / kWave];
  extern __shared__ __align__(16) char s_dyn_raw[];
  // H_c F_c of each direction of the batch, re words then im words: split
  // [dir_batch][n_omega][NW] (one record per omega, 16 bytes with two words,
  // read in one access); FP64 [dir_batch][NW][n_omega] (word-major keeps the
  // FP64 sweep within 128 registers)
  Acc *const s_cf = (Acc *)s_dyn_raw;
  // FP64 sin/cos table (sincos_table)
  constexpr bool kTab = !kSplit && (SR_F64_TABLE != 0);
  __shared__ double2 s_sct[kTab ? kSinCosTab : 1];
  __shared__ double s_wmax[kTab ? kMaxT / kWave : 1];
  // and the largest |omega| of the grid: a step whose |tau| * om_abs_max
  // reaches kTabMaxPhase / 2 takes the table's wide-phase form
  double om_abs_max = 0.0;
  if (kTab) {
    for (int i = threadIdx.x; i < kSinCosTab; i += blockDim.x) {
      double sv, cv;
      sincospi((double)i / (kSinCosTab / 2), &sv, &cv);
      s_sct[i] = make_double2(sv, cv);
    }
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < (int64_t)p.n_omega; i += blockDim.x) m = fmax(m, fabs(p.omega[i]));
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (threadIdx.x % kWave == 0) s_wmax[threadIdx.x / kWave] = m;
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x / kWave); w++) om_abs_max = fmax(om_abs_max, s_wmax[w]);
  }

  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int64_t lp = (int64_t)blockIdx.x % p.chunk_n;
  const int q = (int)((int64_t)blockIdx.x / p.chunk_n);
  const int64_t ip = p.chunk_base + lp;
  const double dt = p.dt;
  // the omega tile [w0, w0 + n_om) whose tables are in LDS
  int w0 = 0, n_om = 0;
  auto cf_at = [&](int d, int j, int k) -> size_t {
    return kSplit ? ((size_t)d * n_om + j) * NW + k : ((size_t)d * NW + k) * n_om + j;
  };

  const Particle pt = load_particle(p, ip);
  double gw_acc = 0.0;
  int64_t n_use = 0;
  if (pt.n > 0) {
    SnapCursor cur;
    cur.init(pt, 1);
    const int64_t n_act = min(pt.n - 1, (int64_t)pt.ite - 1);
    if (cur.next >= 0) n_use = min(n_act, cur.next + 1);
  }
  const double wdt2 = pt.w * dt * dt;
  const uint32_t d_lo = (uint32_t)((int64_t)p.n_dir * q / p.n_groups);
  const uint32_t d_hi = (uint32_t)((int64_t)p.n_dir * (q + 1) / p.n_groups);
  double *slab = p.slab + (int64_t)q * kSlabWords * p.n_points;
  const int64_t fstride = p.chunk_n * p.n_grid;

  // geometry of direction dir: n, e_theta, e_phi
  auto geom = [&](uint32_t dir, double *n, double *et, double *ep) {
    const uint32_t iPhi = dir / p.n_theta;
    const uint32_t iTh = dir - iPhi * p.n_theta;
    const double st = p.sin_t[iTh], ct = p.cos_t[iTh], sp = p.sin_p[iPhi], cp = p.cos_p[iPhi];
    n[0] = st * cp; n[1] = st * sp; n[2] = ct;
    et[0] = ct * cp; et[1] = ct * sp; et[2] = -st;
    ep[0] = -sp; ep[1] = cp; ep[2] = 0.0;
  };

  // directions are taken in batches whose coefficient tables fit the block's
  // LDS, so that each step's nine gradient words are summed over a batch in
  // registers and written once per batch
  for (uint32_t d0 = d_lo; d0 < d_hi && n_use > 0; d0 += (uint32_t)p.dir_batch)
  for (w0 = 0; w0 < (int)p.n_omega; w0 += kAdjOmegaTile) {
    const int nb = (int)min((uint32_t)p.dir_batch, d_hi - d0);
    n_om = min(kAdjOmegaTile, (int)p.n_omega - w0);

    // ---- coefficient tables from the particle's fields and dL/dS ----
    __syncthreads();   // the previous batch is done with the tables
    for (int idx = tid; idx < nb * n_om; idx += T) {
      const int d = idx / n_om;
      const int wi = idx - d * n_om;      // within the tile
      const uint32_t dir = d0 + (uint32_t)d;
      double nv[3], et[3], ep[3];
      geom(dir, nv, et, ep);
      const int64_t gidx = (int64_t)dir * p.n_omega + w0 + wi;
      const int64_t fi = lp * p.n_grid + gidx;
      double Fr[3], Fi[3], G[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Fr[c] = p.F[(int64_t)(2 * c) * fstride + fi];
        Fi[c] = p.F[(int64_t)(2 * c + 1) * fstride + fi];
        G[c] = p.dS[p.comp == SR_COMP_TOTAL ? 0 : c][gidx];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) gw_acc += dt * dt * G[c] * (Fr[c] * Fr[c] + Fi[c] * Fi[c]);
      if constexpr (TWO) {
        const double H = 2.0 * wdt2 * G[0];
        s_cf[cf_at(d, wi, 0)] = (Acc)(H * (Fr[0] * et[0] + Fr[1] * et[1] + Fr[2] * et[2]));
        s_cf[cf_at(d, wi, 1)] = (Acc)(H * (Fr[0] * ep[0] + Fr[1] * ep[1]));
        s_cf[cf_at(d, wi, 2)] = (Acc)(H * (Fi[0] * et[0] + Fi[1] * et[1] + Fi[2] * et[2]));
        s_cf[cf_at(d, wi, 3)] = (Acc)(H * (Fi[0] * ep[0] + Fi[1] * ep[1]));
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double H = 2.0 * wdt2 * G[c];
          s_cf[cf_at(d, wi, c)] = (Acc)(H * Fr[c]);
          s_cf[cf_at(d, wi, 3 + c)] = (Acc)(H * Fi[c]);
        }
      }
      if (d == 0) {
        s_cw[wi] = kSplit ? p.omega[w0 + wi] * kInv2Pi : p.omega[w0 + wi];
        if (kSplit) s_cwf[wi] = (float)s_cw[wi];
      }
    }
    __syncthreads();   // tables complete
    const bool rot = kSplit && p.domega_rev > 0.0 && s_cw[0] >= 0.0;

    // ---- transpose sweep: threads = steps, loop over directions and omegas ----
    for (int64_t t0 = 0; t0 < n_use; t0 += T) {
      const int64_t s = t0 + tid;
      if (s >= n_use) continue;
      const int64_t gs = pt.off + s;
      const double time = (double)(uint32_t)(pt.its + (uint32_t)s) * dt;
      const double xs = p.x[gs], ys = p.y[gs], zs = p.z[gs];
      double tp = 0.0, xp = 0.0, yp = 0.0, zp = 0.0;
      if (s > 0) {
        tp = (double)(uint32_t)(pt.its + (uint32_t)(s - 1)) * dt;
        xp = p.x[gs - 1]; yp = p.y[gs - 1]; zp = p.z[gs - 1];
      }
      // direction-independent part of the step (kernel_farfield.cl:74-83)
      const double *rec = p.pre + 6 * gs;
      const double a[3] = {rec[0], rec[1], rec[2]};
      const double m[3] = {rec[3], rec[4], rec[5]};
      double acc9[kSlabWords];
#pragma unroll
      for (int k = 0; k < kSlabWords; k++) acc9[k] = 0.0;

      for (int d = 0; d < nb; d++) {
        double nv[3], et[3], ep[3];
        geom(d0 + (uint32_t)d, nv, et, ep);

        const double tau = time - (xs * nv[0] + ys * nv[1] + zs * nv[2]);
        const double tau_prev = (s > 0) ? tp - (xp * nv[0] + yp * nv[1] + zp * nv[2]) : 0.0;
        double c1 = a[0] * nv[0] + a[1] * nv[1] + a[2] * nv[2];
        double c2 = 1.0 - (m[0] * nv[0] + m[1] * nv[1] + m[2] * nv[2]);
        c2 = 1.0 / c2;
        c1 = c1 * c2 * c2;
        const double A[3] = {c1 * (nv[0] - m[0]) - c2 * a[0], c1 * (nv[1] - m[1]) - c2 * a[1],
                             c1 * (nv[2] - m[2]) - c2 * a[2]};

        Acc P[NCW], R[NCW];
#pragma unroll
        for (int c = 0; c < NCW; c++) { P[c] = 0; R[c] = 0; }
        auto acc_pair = [&](Acc sn, Acc cs, int j) {
          const Acc wr = kSplit ? (Acc)s_cwf[kSplit ? j : 0] : (Acc)s_cw[j];
          if constexpr (TWO && kSplit) {
            const float4 v = *(const float4 *)&s_cf[cf_at(d, j, 0)];
            const float ar[2] = {v.x, v.y}, ai[2] = {v.z, v.w};
#pragma unroll
            for (int c = 0; c < 2; c++) {
              P[c] = fma(ar[c], cs, fma(ai[c], sn, P[c]));
              R[c] = fma(wr, fma(ar[c], sn, -ai[c] * cs), R[c]);
            }
          } else {
#pragma unroll
            for (int c = 0; c < NCW; c++) {
              const Acc ar = s_cf[cf_at(d, j, c)], ai = s_cf[cf_at(d, j, NCW + c)];
              P[c] = fma(ar, cs, fma(ai, sn, P[c]));
              R[c] = fma(wr, fma(ar, sn, -ai * cs), R[c]);
            }
          }
        };
        if (rot) {
          // the omegas this step keeps: those below its threshold (a prefix)
          const double thr = 0.5 / fabs(tau - tau_prev);
          int lo = 0, hi = n_om;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_cw[mid] < thr) lo = mid + 1; else hi = mid;
          }
          const int jmax = lo;
          float er, ei, sn = 0.f, cn = 1.f;
          sincos_rev(p.domega_rev * tau, &ei, &er);
#pragma unroll SR_ADJ_ROT_UNROLL
          for (int j = 0; j < jmax; j++) {
            if ((j & 31) == 0) {
              sincos_rev(s_cw[j] * tau, &sn, &cn);
            } else {
              const float cn2 = __builtin_fmaf(cn, er, -sn * ei);
              sn = __builtin_fmaf(cn, ei, sn * er);
              cn = cn2;
            }
            acc_pair((Acc)sn, (Acc)cn, j);
          }
        } else {
          const double thr = kSplit ? 0.5 / fabs(tau - tau_prev) : 0.0;
          // WIDE: a step of the wave may reach the sin/cos table's limit (a
          // compile-time flag, so the usual loop is the table form alone)
          auto om_sweep = [&](auto Wc) {
#pragma unroll SR_ADJ_UNROLL
            for (int j = 0; j < n_om; j++) {
              const double w = s_cw[j];
              const double ph = w * tau;
              if (kSplit) {
                if (!(fabs(w) < thr)) continue;   // the forward's threshold form
              } else {
                const double php = (s > 0) ? w * tau_prev : 0.0;
                if (!(fabs(ph - php) < kPi)) continue;   // kernel_farfield.cl:69-72, exactly
              }
              Acc sn, cs;
              if (kSplit) {
                float s32, c32;
                sincos_rev(ph, &s32, &c32);
                sn = s32; cs = c32;
              } else {
                double s64, c64;
                if (kTab)
                  sincos_table(ph, s_sct, decltype(Wc)::value != 0, &s64, &c64);
                else
                  sincos_reduced(ph, &s64, &c64);
                sn = s64; cs = c64;
              }
              acc_pair(sn, cs, j);
            }
          };
          if (kTab && __ballot(fabs(tau) * om_abs_max >= 0.5 * kTabMaxPhase) != 0)
            om_sweep(std::integral_constant<int, 1>());
          else
            om_sweep(std::integral_constant<int, 0>());
        }
        double gA[3], gR[3];
        if constexpr (TWO) {
#pragma unroll
          for (int c = 0; c < 3; c++) {
            gA[c] = (double)P[0] * et[c] + (double)P[1] * ep[c];
            gR[c] = (double)R[0] * et[c] + (double)R[1] * ep[c];
          }
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) { gA[c] = (double)P[c]; gR[c] = (double)R[c]; }
        }
        const double rs = kSplit ? 2.0 * kPi : 1.0;   // revolutions -> radians
        const double AR = rs * (A[0] * gR[0] + A[1] * gR[1] + A[2] * gR[2]);
        const double nmg = (nv[0] - m[0]) * gA[0] + (nv[1] - m[1]) * gA[1] + (nv[2] - m[2]) * gA[2];
        const double ag = a[0] * gA[0] + a[1] * gA[1] + a[2] * gA[2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          acc9[c] += nv[c] * AR;
          acc9[3 + c] += 2.0 * c1 * c2 * nmg * nv[c] - c1 * gA[c] - c2 * c2 * ag * nv[c];
          acc9[6 + c] += c2 * c2 * nmg * nv[c] - c2 * gA[c];
        }
      }
#pragma unroll
      for (int k = 0; k < kSlabWords; k++) slab[(int64_t)k * p.n_points + gs] += acc9[k];
    }
  }

  // weight gradient of this (particle, group): block sum in a fixed order
  double v = gw_acc;
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  if ((tid & (kWave - 1)) == 0) s_red[tid / kWave] = v;
  __syncthreads();
  if (tid == 0) {
    double sum = 0.0;
    for (int w = 0; w < T / kWave; w++) sum += s_red[w];
    p.slab_w[(int64_t)q * p.n_particles + ip] = sum;
  }
}

// sum the group slabs of one sample and chain dL/dm, dL/da -> dL/db -> dL/du
__global__ void k_adjoint_reduce(const AParams p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n_points) return;
  // particle of sample i: last particle whose first sample is <= i
  int64_t lo = 0, hi = p.n_particles;   // offsets[lo] <= i < offsets[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (p.offsets[mid] <= i) lo = mid; else hi = mid;
  }
  const bool has_prev = i > p.offsets[lo];
  double g[kSlabWords], gp[6];
  for (int k = 0; k < kSlabWords; k++) g[k] = 0.0;
  for (int k = 0; k < 6; k++) gp[k] = 0.0;
  for (int q = 0; q < p.n_groups; q++) {
    const double *sl = p.slab + (int64_t)q * kSlabWords * p.n_points;
    for (int k = 0; k < kSlabWords; k++) g[k] += sl[(int64_t)k * p.n_points + i];
    if (has_prev)
      for (int k = 0; k < 6; k++) gp[k] += sl[(int64_t)(3 + k) * p.n_points + i - 1];
  }
  const double dtInv = 1.0 / p.dt;
  double gb[3];
  for (int c = 0; c < 3; c++) gb[c] = 0.5 * g[3 + c] - dtInv * g[6 + c] + 0.5 * gp[c] + dtInv * gp[3 + c];
  const double ux = p.ux[i], uy = p.uy[i], uz = p.uz[i];
  const double gi = gamma_inv(ux, uy, uz);
  const double bx = ux * gi, by = uy * gi, bz = uz * gi;
  const double bg = bx * gb[0] + by * gb[1] + bz * gb[2];
  p.gx[i] += g[0];
  p.gy[i] += g[1];
  p.gz[i] += g[2];
  p.gux[i] += gi * (gb[0] - bx * bg);
  p.guy[i] += gi * (gb[1] - by * bg);
  p.guz[i] += gi * (gb[2] - bz * bg);
}

__global__ void k_adjoint_reduce_w(const AParams p) {
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= p.n_particles) return;
  double s = 0.0;
  for (int q = 0; q < p.n_groups; q++) s += p.slab_w[(int64_t)q * p.n_particles + ip];
  p.gw[ip] += s;
}

__global__ void k_adjoint_clear(double *buf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[i] = 0.0;
}

// Spectra from the per-particle fields F[6][n_particles][n_grid] (x, y, z
// as re/im pairs): out += w dt^2 |F|^2 particle by particle, in particle order
// and with the record's own expressions (spectrum.hip emit_k), so the sums
// equal those of the spectrum kernel run without particle slices.
__global__ void k_fields_spectrum(const int64_t *offsets, const double *weight, double dt, int64_t n_particles,
                                  int64_t n_grid, int32_t comp, const double *F, double *out0, double *out1,
                                  double *out2) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_grid) return;
  const int64_t fs = n_particles * n_grid;
  double acc[3] = {out0[g], comp == SR_COMP_TOTAL ? 0.0 : out1[g], comp == SR_COMP_TOTAL ? 0.0 : out2[g]};
  for (int64_t ip = 0; ip < n_particles; ip++) {
    if (offsets[ip + 1] <= offsets[ip]) continue;   // no samples: no record
    const double wpdt2 = weight[ip] * dt * dt;
    const double *f = F + ip * n_grid + g;
    double re[3], im[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      re[c] = f[(2 * c) * fs];
      im[c] = f[(2 * c + 1) * fs];
    }
    if (comp == SR_COMP_TOTAL) {
      acc[0] += wpdt2 * ((re[0] * re[0] + re[1] * re[1] + re[2] * re[2]) +
                         (im[0] * im[0] + im[1] * im[1] + im[2] * im[2]));
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] += wpdt2 * (re[c] * re[c] + im[c] * im[c]);
    }
  }
  out0[g] = acc[0];
  if (comp != SR_COMP_TOTAL) {
    out1[g] = acc[1];
    out2[g] = acc[2];
  }
}

struct APlan {
  int threads;
  int groups;
  int64_t chunk;   // particles whose fields share the workspace at once
  int dir_batch;   // directions per LDS batch
  size_t lds;      // dynamic LDS of the sweep kernel
};

// LDS for the coefficient tables of one direction batch
constexpr size_t kAdjTableBudget = 32768;

// fields of one chunk of particles: 6 doubles per (particle, point), ~2 GiB at most
constexpr double kFieldBudget = 2147483648.0;

// all_fields: the caller holds every particle's fields (one chunk)
APlan adjoint_plan(const sr_tracks &t, const sr_grid &g, const sr_opts &o, bool all_fields) {
  APlan pl;
  // threads = steps of the transpose sweep
  const int waves = o.waves_per_block > 0 ? o.waves_per_block : 4;
  pl.threads = kWave * waves;
  const int64_t n_dir = (int64_t)g.n_theta * g.n_phi;
  const int64_t n_grid = n_dir * g.n_omega;
  pl.chunk = all_fields ? std::max<int64_t>(1, t.n_particles)
                        : std::max<int64_t>(1, std::min<int64_t>(t.n_particles,
                                                                 (int64_t)(kFieldBudget / (48.0 * n_grid))));
  int64_t groups = o.slices > 0 ? o.slices : std::max<int64_t>(1, (4096 + pl.chunk - 1) / pl.chunk);
  pl.groups = (int)std::min<int64_t>(std::min<int64_t>(groups, n_dir), 1 << 16);
  const bool split = o.precision == SR_PREC_SPLIT;
  const size_t per_dir = (size_t)((o.comp == SR_COMP_TOTAL) ? 4 : 6) * std::min<uint32_t>(g.n_omega, kAdjOmegaTile) *
                         (split ? sizeof(float) : sizeof(double));
  const int64_t dirs_per_group = (n_dir + pl.groups - 1) / pl.groups;
  pl.dir_batch = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)(kAdjTableBudget / per_dir), 16,
                                                              dirs_per_group}));
  pl.lds = per_dir * pl.dir_batch;
  return pl;
}

int adjoint_check(const sr_tracks *t, const sr_grid *g, const sr_opts *o) {
  if (!t || !g || !o) return fail(SR_ERR_INVALID, "tracks, grid and opts are required");
  if (g->mode != SR_MODE_FAR || g->screen_y) return fail(SR_ERR_UNSUPPORTED, "the adjoint covers the far field only");
  if (o->comp != SR_COMP_TOTAL && o->comp != SR_COMP_CARTESIAN)
    return fail(SR_ERR_UNSUPPORTED, "the adjoint covers comp 'total' and 'cartesian' only");
  if (o->precision != SR_PREC_F64 && o->precision != SR_PREC_SPLIT)
    return fail(SR_ERR_INVALID, "precision %d is not F64 (0) or SPLIT (1)", o->precision);
  if (g->n_omega == 0 || g->n_theta == 0 || g->n_phi == 0)
    return fail(SR_ERR_INVALID, "grid sizes must be >= 1");
  if (!g->omega || !g->sin_theta || !g->cos_theta || !g->sin_phi || !g->cos_phi)
    return fail(SR_ERR_INVALID, "grid axis pointer is NULL");
  if (t->n_particles < 0 || t->n_points < 0) return fail(SR_ERR_INVALID, "negative track sizes");
  if (t->n_particles > 0) {
    if (!t->offsets || !t->x || !t->y || !t->z || !t->ux || !t->uy || !t->uz || !t->weight ||
        !t->it_start || !t->it_end || !t->it_snaps)
      return fail(SR_ERR_INVALID, "a track pointer is NULL");
    if (t->n_snaps != 1) return fail(SR_ERR_UNSUPPORTED, "the adjoint needs exactly one snapshot");
    if (t->snaps_stride > 1) return fail(SR_ERR_INVALID, "snaps_stride must be 0 or 1");
    if (!(t->dt != 0.0) || !isfinite(t->dt)) return fail(SR_ERR_INVALID, "dt must be finite and non-zero");
  }
  if (o->waves_per_block < 0 || o->waves_per_block > 4 || o->waves_per_block == 3)
    return fail(SR_ERR_INVALID, "waves_per_block must be 0, 1, 2 or 4");
  if (o->slices < 0) return fail(SR_ERR_INVALID, "slices (direction groups) must be >= 0");
  return SR_OK;
}

size_t adjoint_slab_words(const sr_tracks &t, const APlan &pl) {
  return (size_t)pl.groups * kSlabWords * (size_t)t.n_points + (size_t)pl.groups * (size_t)t.n_particles;
}

// workspace: slabs, the fields of one chunk of particles (unless the caller
// passes them), the step data, and (with own fields) the fields launches' tile
// order
size_t adjoint_ws(const sr_tracks &t, const sr_grid &g, const APlan &pl, bool own_fields, int32_t precision) {
  const size_t n_grid = (size_t)g.n_omega * g.n_theta * g.n_phi;
  return (adjoint_slab_words(t, pl) + (own_fields ? 6 * (size_t)pl.chunk * n_grid : 0) +
          6 * (size_t)t.n_points) * sizeof(double) +
         (own_fields ? fields_order_space(g, precision, pl.chunk) : 0);
}

// the adjoint; fields_in = NULL: the fields are computed chunk by chunk in the
// workspace, else they are the caller's F[6][n_particles][n_grid]
int adjoint_run(const sr_tracks *t, const sr_grid *g, const sr_opts *o, const double *const *dS,
                const double *fields_in, double *gx, double *gy, double *gz, double *gux, double *guy,
                double *guz, double *gw, void *workspace, size_t workspace_bytes, void *stream) {
  int rc = adjoint_check(t, g, o);
  if (rc) return rc;
  if (t->n_particles == 0) return SR_OK;
  const int n_in = (o->comp == SR_COMP_TOTAL) ? 1 : 3;
  if (!dS) return fail(SR_ERR_INVALID, "dS is NULL");
  for (int c = 0; c < n_in; c++)
    if (!dS[c]) return fail(SR_ERR_INVALID, "dS[%d] is NULL", c);
  if (!gx || !gy || !gz || !gux || !guy || !guz || !gw) return fail(SR_ERR_INVALID, "a gradient pointer is NULL");
  const bool own = (fields_in == nullptr);
  const APlan pl = adjoint_plan(*t, *g, *o, !own);
  const size_t need = adjoint_ws(*t, *g, pl, own, o->precision);
  if (!workspace || workspace_bytes < need)
    return fail(SR_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  AParams p;
  memset(&p, 0, sizeof(p));
  p.offsets = t->offsets;
  p.x = t->x; p.y = t->y; p.z = t->z; p.ux = t->ux; p.uy = t->uy; p.uz = t->uz;
  p.weight = t->weight;
  p.it_start = t->it_start; p.it_end = t->it_end; p.it_snaps = t->it_snaps;
  p.n_snaps = 1;
  p.snaps_stride = t->snaps_stride;
  p.dt = t->dt;
  p.n_particles = t->n_particles;
  p.n_points = t->n_points;
  p.comp = o->comp;
  p.n_omega = g->n_omega; p.n_theta = g->n_theta; p.n_phi = g->n_phi;
  p.n_dir = g->n_theta * g->n_phi;
  p.omega = g->omega; p.sin_t = g->sin_theta; p.cos_t = g->cos_theta; p.sin_p = g->sin_phi; p.cos_p = g->cos_phi;
  p.domega_rev = g->omega_step * kInv2Pi;
  p.n_grid = (int64_t)g->n_omega * g->n_theta * g->n_phi;
  for (int c = 0; c < 3; c++) p.dS[c] = dS[c < n_in ? c : 0];
  p.n_groups = pl.groups;
  p.dir_batch = pl.dir_batch;
  p.slab = (double *)workspace;
  p.slab_w = p.slab + (size_t)pl.groups * kSlabWords * (size_t)t->n_points;
  p.gx = gx; p.gy = gy; p.gz = gz; p.gux = gux; p.guy = guy; p.guz = guz; p.gw = gw;
  double *fields = (double *)workspace + adjoint_slab_words(*t, pl);
  double *pre = own ? fields + 6 * (size_t)pl.chunk * (size_t)p.n_grid : fields;
  void *order_ws = own ? (void *)(pre + 6 * (size_t)t->n_points) : nullptr;
  const size_t order_space = own ? fields_order_space(*g, o->precision, pl.chunk) : 0;
  p.F = own ? fields : fields_in;
  p.pre = pre;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n_clear = (int64_t)adjoint_slab_words(*t, pl);
  hipLaunchKernelGGL(k_adjoint_clear, dim3((unsigned)((n_clear + 255) / 256)), dim3(256), 0, st, p.slab, n_clear);
  if ((rc = step_prep(*t, SR_MODE_FAR, pre, st))) return rc;
  for (int64_t pb = 0; pb < t->n_particles; pb += pl.chunk) {
    const int64_t pe = std::min<int64_t>(t->n_particles, pb + pl.chunk);
    if (own) {
      const int64_t n_f = 6 * (pe - pb) * p.n_grid;
      hipLaunchKernelGGL(k_adjoint_clear, dim3((unsigned)((n_f + 255) / 256)), dim3(256), 0, st, fields, n_f);
      if ((rc = spectrum_fields(*t, *g, o->precision, pb, pe, pre, fields, st, order_ws, order_space))) return rc;
    }
    p.chunk_base = pb;
    p.chunk_n = pe - pb;
    const int64_t blocks = p.chunk_n * (int64_t)pl.groups;
    if (o->precision == SR_PREC_F64 && o->comp == SR_COMP_TOTAL)
      hipLaunchKernelGGL((k_adjoint<SR_PREC_F64, true>), dim3((unsigned)blocks), dim3(pl.threads), pl.lds, st, p);
    else if (o->precision == SR_PREC_F64)
      hipLaunchKernelGGL((k_adjoint<SR_PREC_F64, false>), dim3((unsigned)blocks), dim3(pl.threads), pl.lds, st, p);
    else if (o->comp == SR_COMP_TOTAL)
      hipLaunchKernelGGL((k_adjoint<SR_PREC_SPLIT, true>), dim3((unsigned)blocks), dim3(pl.threads), pl.lds, st, p);
    else
      hipLaunchKernelGGL((k_adjoint<SR_PREC_SPLIT, false>), dim3((unsigned)blocks), dim3(pl.threads), pl.lds, st, p);
  }
  hipLaunchKernelGGL(k_adjoint_reduce, dim3((unsigned)((t->n_points + 255) / 256)), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_adjoint_reduce_w, dim3((unsigned)((t->n_particles + 255) / 256)), dim3(256), 0, st, p);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SR_ERR_HIP, "adjoint launch: %s", hipGetErrorString(e));
  return SR_OK;
}

}  // namespace

extern "C" {

int sr_far_adjoint_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                                   size_t *bytes) {
  if (!bytes) return fail(SR_ERR_INVALID, "bytes is NULL");
  *bytes = 0;
  int rc = adjoint_check(tracks, grid, opts);
  if (rc) return rc;
  if (tracks->n_particles == 0) return SR_OK;
  *bytes = adjoint_ws(*tracks, *grid, adjoint_plan(*tracks, *grid, *opts, false), true, opts->precision);
  return SR_OK;
}

int sr_far_adjoint_batch(const sr_tracks *t, const sr_grid *g, const sr_opts *o, const double *const *dS,
                         double *gx, double *gy, double *gz, double *gux, double *guy, double *guz,
                         double *gw, void *workspace, size_t workspace_bytes, void *stream) {
  return adjoint_run(t, g, o, dS, nullptr, gx, gy, gz, gux, guy, guz, gw, workspace, workspace_bytes, stream);
}

int sr_far_fields_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                                  size_t *bytes) {
  if (!bytes) return fail(SR_ERR_INVALID, "bytes is NULL");
  *bytes = 0;
  int rc = adjoint_check(tracks, grid, opts);
  if (rc) return rc;
  *bytes = 6 * (size_t)tracks->n_points * sizeof(double) +
           fields_order_space(*grid, opts->precision, tracks->n_particles);
  return SR_OK;
}

int sr_far_fields_batch(const sr_tracks *t, const sr_grid *g, const sr_opts *o, double *fields,
                        double *const *outs, void *workspace, size_t workspace_bytes, void *stream) {
  int rc = adjoint_check(t, g, o);
  if (rc) return rc;
  if (t->n_particles == 0) return SR_OK;
  if (!fields) return fail(SR_ERR_INVALID, "fields is NULL");
  const int n_out = (o->comp == SR_COMP_TOTAL) ? 1 : 3;
  if (outs)
    for (int c = 0; c < n_out; c++)
      if (!outs[c]) return fail(SR_ERR_INVALID, "outs[%d] is NULL", c);
  const size_t order_space = fields_order_space(*g, o->precision, t->n_particles);
  const size_t need = 6 * (size_t)t->n_points * sizeof(double) + order_space;
  if (!workspace || workspace_bytes < need)
    return fail(SR_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  hipStream_t st = (hipStream_t)stream;
  double *pre = (double *)workspace;
  const int64_t n_grid = (int64_t)g->n_omega * g->n_theta * g->n_phi;
  const int64_t n_f = 6 * t->n_particles * n_grid;
  hipLaunchKernelGGL(k_adjoint_clear, dim3((unsigned)((n_f + 255) / 256)), dim3(256), 0, st, fields, n_f);
  if ((rc = step_prep(*t, SR_MODE_FAR, pre, st))) return rc;
  if ((rc = spectrum_fields(*t, *g, o->precision, 0, t->n_particles, pre, fields, st,
                            (void *)(pre + 6 * (size_t)t->n_points), order_space)))
    return rc;
  if (outs)
    hipLaunchKernelGGL(k_fields_spectrum, dim3((unsigned)((n_grid + 255) / 256)), dim3(256), 0, st, t->offsets,
                       t->weight, t->dt, t->n_particles, n_grid, o->comp, (const double *)fields, outs[0],
                       n_out > 1 ? outs[1] : nullptr, n_out > 1 ? outs[2] : nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SR_ERR_HIP, "fields launch: %s", hipGetErrorString(e));
  return SR_OK;
}

int sr_far_adjoint_fields_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                                          size_t *bytes) {
  if (!bytes) return fail(SR_ERR_INVALID, "bytes is NULL");
  *bytes = 0;
  int rc = adjoint_check(tracks, grid, opts);
  if (rc) return rc;
  if (tracks->n_particles == 0) return SR_OK;
  *bytes = adjoint_ws(*tracks, *grid, adjoint_plan(*tracks, *grid, *opts, true), false, opts->precision);
  return SR_OK;
}

int sr_far_adjoint_from_fields(const sr_tracks *t, const sr_grid *g, const sr_opts *o, const double *const *dS,
                               const double *fields, double *gx, double *gy, double *gz, double *gux,
                               double *guy, double *guz, double *gw, void *workspace, size_t workspace_bytes,
                               void *stream) {
  if (t && t->n_particles > 0 && !fields) return fail(SR_ERR_INVALID, "fields is NULL");
  return adjoint_run(t, g, o, dS, fields, gx, gy, gz, gux, guy, guz, gw, workspace, workspace_bytes, stream);
}

}  // extern "C"
