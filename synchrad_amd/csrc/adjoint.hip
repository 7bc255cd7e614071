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

constexpr int kAdjMaxOmega = 512;  // per-omega coefficient table held in LDS
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
  int64_t chunk_base, chunk_n, n_grid;
  // decomposition and workspace
  int32_t n_groups;
  double *slab;    // [n_groups][kSlabWords][n_points]
  double *slab_w;  // [n_groups][n_particles]
  // outputs (+=)
  double *gx, *gy, *gz, *gux, *guy, *guz, *gw;
};

// TWO: split 'total', coefficients on e_theta (word 0) and e_phi (word 1)
template <int PREC, bool TWO>
__global__ void __launch_bounds__(kMaxT) k_adjoint(const AParams p) {
  constexpr bool kSplit = (PREC == SR_PREC_SPLIT);
  static_assert(kSplit || !TWO, "two-word coefficients are split-only");
  using Acc = typename std::conditional<kSplit, float, double>::type;
  __shared__ double s_cw[kAdjMaxOmega];     // omega (F64) or omega/2pi (split)
  __shared__ Acc s_cf[6][kAdjMaxOmega];     // H_c Re F_c (0..2), H_c Im F_c (3..5)
  __shared__ double s_red[kMaxT / kWave];

  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int64_t lp = (int64_t)blockIdx.x % p.chunk_n;
  const int q = (int)((int64_t)blockIdx.x / p.chunk_n);
  const int64_t ip = p.chunk_base + lp;
  const double dt = p.dt;
  const double dtInv = 1.0 / dt;
  constexpr bool two = TWO;
  constexpr int NCW = TWO ? 2 : 3;   // coefficient words per omega (re and im each)

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

  for (uint32_t dir = d_lo; dir < d_hi && n_use > 0; dir++) {
    const uint32_t iPhi = dir / p.n_theta;
    const uint32_t iTh = dir - iPhi * p.n_theta;
    const double st = p.sin_t[iTh], ct = p.cos_t[iTh], sp = p.sin_p[iPhi], cp = p.cos_p[iPhi];
    const double nx = st * cp, ny = st * sp, nz = ct;
    const double et[3] = {ct * cp, ct * sp, -st};   // e_theta
    const double ep[3] = {-sp, cp, 0.0};            // e_phi

    // ---- coefficient table from the particle's field and dL/dS ----
    __syncthreads();   // the previous direction is done with the table
    for (int wi = tid; wi < (int)p.n_omega; wi += T) {
      const int64_t gidx = (int64_t)dir * p.n_omega + wi;
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
        s_cf[0][wi] = (Acc)(H * (Fr[0] * et[0] + Fr[1] * et[1] + Fr[2] * et[2]));
        s_cf[1][wi] = (Acc)(H * (Fr[0] * ep[0] + Fr[1] * ep[1]));
        s_cf[3][wi] = (Acc)(H * (Fi[0] * et[0] + Fi[1] * et[1] + Fi[2] * et[2]));
        s_cf[4][wi] = (Acc)(H * (Fi[0] * ep[0] + Fi[1] * ep[1]));
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const double H = 2.0 * wdt2 * G[c];
          s_cf[c][wi] = (Acc)(H * Fr[c]);
          s_cf[3 + c][wi] = (Acc)(H * Fi[c]);
        }
      }
      s_cw[wi] = kSplit ? p.omega[wi] * kInv2Pi : p.omega[wi];
    }
    __syncthreads();   // table complete
    const bool rot = kSplit && p.domega_rev > 0.0 && s_cw[0] >= 0.0;

    // ---- transpose sweep: threads = steps, loop over the omega table ----
    for (int64_t t0 = 0; t0 < n_use; t0 += T) {
      const int64_t s = t0 + tid;
      if (s >= n_use) continue;
      const int64_t gs = pt.off + s;
      const double time = (double)(uint32_t)(pt.its + (uint32_t)s) * dt;
      const double tau = time - (p.x[gs] * nx + p.y[gs] * ny + p.z[gs] * nz);
      double tau_prev = 0.0;
      if (s > 0) {
        const double tp = (double)(uint32_t)(pt.its + (uint32_t)(s - 1)) * dt;
        tau_prev = tp - (p.x[gs - 1] * nx + p.y[gs - 1] * ny + p.z[gs - 1] * nz);
      }
      const double u0x = p.ux[gs], u0y = p.uy[gs], u0z = p.uz[gs];
      const double u1x = p.ux[gs + 1], u1y = p.uy[gs + 1], u1z = p.uz[gs + 1];
      const double g0 = gamma_inv(u0x, u0y, u0z), g1 = gamma_inv(u1x, u1y, u1z);
      const double b0x = u0x * g0, b0y = u0y * g0, b0z = u0z * g0;
      const double b1x = u1x * g1, b1y = u1y * g1, b1z = u1z * g1;
      const double a[3] = {(b1x - b0x) * dtInv, (b1y - b0y) * dtInv, (b1z - b0z) * dtInv};
      const double m[3] = {0.5 * (b1x + b0x), 0.5 * (b1y + b0y), 0.5 * (b1z + b0z)};
      double c1 = a[0] * nx + a[1] * ny + a[2] * nz;
      double c2 = 1.0 - (m[0] * nx + m[1] * ny + m[2] * nz);
      c2 = 1.0 / c2;
      c1 = c1 * c2 * c2;
      const double A[3] = {c1 * (nx - m[0]) - c2 * a[0], c1 * (ny - m[1]) - c2 * a[1],
                           c1 * (nz - m[2]) - c2 * a[2]};

      Acc P[3] = {0, 0, 0}, R[3] = {0, 0, 0};
      auto acc_pair = [&](Acc sn, Acc cs, int j) {
        const Acc wr = (Acc)s_cw[j];
#pragma unroll
        for (int c = 0; c < NCW; c++) {
          const Acc ar = s_cf[c][j], ai = s_cf[3 + c][j];
          P[c] = fma(ar, cs, fma(ai, sn, P[c]));
          R[c] = fma(wr, fma(ar, sn, -ai * cs), R[c]);
        }
      };
      if (rot) {
        // the omegas this step keeps: those below its threshold (a prefix)
        const double thr = 0.5 / fabs(tau - tau_prev);
        int lo = 0, hi = (int)p.n_omega;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_cw[mid] < thr) lo = mid + 1; else hi = mid;
        }
        const int jmax = lo;
        float er, ei, sn = 0.f, cn = 1.f;
        sincos_rev(p.domega_rev * tau, &ei, &er);
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
        for (int j = 0; j < (int)p.n_omega; j++) {
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
            sincos_reduced(ph, &s64, &c64);
            sn = s64; cs = c64;
          }
          acc_pair(sn, cs, j);
        }
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
      const double nmg = (nx - m[0]) * gA[0] + (ny - m[1]) * gA[1] + (nz - m[2]) * gA[2];
      const double ag = a[0] * gA[0] + a[1] * gA[1] + a[2] * gA[2];
      const double nv[3] = {nx, ny, nz};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        slab[(int64_t)c * p.n_points + gs] += nv[c] * AR;
        slab[(int64_t)(3 + c) * p.n_points + gs] +=
            2.0 * c1 * c2 * nmg * nv[c] - c1 * gA[c] - c2 * c2 * ag * nv[c];
        slab[(int64_t)(6 + c) * p.n_points + gs] += c2 * c2 * nmg * nv[c] - c2 * gA[c];
      }
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

struct APlan {
  int threads;
  int groups;
  int64_t chunk;   // particles whose fields share the workspace at once
};

// fields of one chunk of particles: 6 doubles per (particle, point), ~2 GiB at most
constexpr double kFieldBudget = 2147483648.0;

APlan adjoint_plan(const sr_tracks &t, const sr_grid &g, const sr_opts &o) {
  APlan pl;
  // threads = steps of the transpose sweep
  const int waves = o.waves_per_block > 0 ? o.waves_per_block : 4;
  pl.threads = kWave * waves;
  const int64_t n_dir = (int64_t)g.n_theta * g.n_phi;
  const int64_t n_grid = n_dir * g.n_omega;
  pl.chunk = std::max<int64_t>(1, std::min<int64_t>(t.n_particles, (int64_t)(kFieldBudget / (48.0 * n_grid))));
  int64_t groups = o.slices > 0 ? o.slices : std::max<int64_t>(1, (4096 + pl.chunk - 1) / pl.chunk);
  pl.groups = (int)std::min<int64_t>(std::min<int64_t>(groups, n_dir), 1 << 16);
  return pl;
}

int adjoint_check(const sr_tracks *t, const sr_grid *g, const sr_opts *o) {
  if (!t || !g || !o) return fail(SR_ERR_INVALID, "tracks, grid and opts are required");
  if (g->mode != SR_MODE_FAR) return fail(SR_ERR_UNSUPPORTED, "the adjoint covers the far field only");
  if (o->comp != SR_COMP_TOTAL && o->comp != SR_COMP_CARTESIAN)
    return fail(SR_ERR_UNSUPPORTED, "the adjoint covers comp 'total' and 'cartesian' only");
  if (o->precision != SR_PREC_F64 && o->precision != SR_PREC_SPLIT)
    return fail(SR_ERR_INVALID, "precision %d is not F64 (0) or SPLIT (1)", o->precision);
  if (g->n_omega == 0 || g->n_theta == 0 || g->n_phi == 0)
    return fail(SR_ERR_INVALID, "grid sizes must be >= 1");
  if (g->n_omega > (uint32_t)kAdjMaxOmega)
    return fail(SR_ERR_UNSUPPORTED, "the adjoint holds at most %d omegas", kAdjMaxOmega);
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

size_t adjoint_ws(const sr_tracks &t, const sr_grid &g, const APlan &pl) {
  const size_t n_grid = (size_t)g.n_omega * g.n_theta * g.n_phi;
  return (adjoint_slab_words(t, pl) + 6 * (size_t)pl.chunk * n_grid) * sizeof(double);
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
  *bytes = adjoint_ws(*tracks, *grid, adjoint_plan(*tracks, *grid, *opts));
  return SR_OK;
}

int sr_far_adjoint_batch(const sr_tracks *t, const sr_grid *g, const sr_opts *o, const double *const *dS,
                         double *gx, double *gy, double *gz, double *gux, double *guy, double *guz,
                         double *gw, void *workspace, size_t workspace_bytes, void *stream) {
  int rc = adjoint_check(t, g, o);
  if (rc) return rc;
  if (t->n_particles == 0) return SR_OK;
  const int n_in = (o->comp == SR_COMP_TOTAL) ? 1 : 3;
  if (!dS) return fail(SR_ERR_INVALID, "dS is NULL");
  for (int c = 0; c < n_in; c++)
    if (!dS[c]) return fail(SR_ERR_INVALID, "dS[%d] is NULL", c);
  if (!gx || !gy || !gz || !gux || !guy || !guz || !gw) return fail(SR_ERR_INVALID, "a gradient pointer is NULL");
  const APlan pl = adjoint_plan(*t, *g, *o);
  const size_t need = adjoint_ws(*t, *g, pl);
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
  p.slab = (double *)workspace;
  p.slab_w = p.slab + (size_t)pl.groups * kSlabWords * (size_t)t->n_points;
  p.gx = gx; p.gy = gy; p.gz = gz; p.gux = gux; p.guy = guy; p.guz = guz; p.gw = gw;
  double *fields = (double *)workspace + adjoint_slab_words(*t, pl);
  p.F = fields;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n_clear = (int64_t)adjoint_slab_words(*t, pl);
  hipLaunchKernelGGL(k_adjoint_clear, dim3((unsigned)((n_clear + 255) / 256)), dim3(256), 0, st, p.slab, n_clear);
  for (int64_t pb = 0; pb < t->n_particles; pb += pl.chunk) {
    const int64_t pe = std::min<int64_t>(t->n_particles, pb + pl.chunk);
    const int64_t n_f = 6 * (pe - pb) * p.n_grid;
    hipLaunchKernelGGL(k_adjoint_clear, dim3((unsigned)((n_f + 255) / 256)), dim3(256), 0, st, fields, n_f);
    if ((rc = spectrum_fields(*t, *g, o->precision, pb, pe, fields, st))) return rc;
    p.chunk_base = pb;
    p.chunk_n = pe - pb;
    const int64_t blocks = p.chunk_n * (int64_t)pl.groups;
    if (o->precision == SR_PREC_F64)
      hipLaunchKernelGGL((k_adjoint<SR_PREC_F64, false>), dim3((unsigned)blocks), dim3(pl.threads), 0, st, p);
    else if (o->comp == SR_COMP_TOTAL)
      hipLaunchKernelGGL((k_adjoint<SR_PREC_SPLIT, true>), dim3((unsigned)blocks), dim3(pl.threads), 0, st, p);
    else
      hipLaunchKernelGGL((k_adjoint<SR_PREC_SPLIT, false>), dim3((unsigned)blocks), dim3(pl.threads), 0, st, p);
  }
  hipLaunchKernelGGL(k_adjoint_reduce, dim3((unsigned)((t->n_points + 255) / 256)), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_adjoint_reduce_w, dim3((unsigned)((t->n_particles + 255) / 256)), dim3(256), 0, st, p);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SR_ERR_HIP, "adjoint launch: %s", hipGetErrorString(e));
  return SR_OK;
}

}  // extern "C"
