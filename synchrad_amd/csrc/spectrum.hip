// spectrum.hip -- gfx950 kernels and C ABI for the spectral radiation integral.
//
// What is computed (per spectral point g and particle p, reference semantics):
//   far : F_p(g) = sum_s A_s(n) exp(i omega (t_s - n.x_s))             kernel_farfield.cl:59-97
//         A_s(n) = c1 (n - bbar_s) - c2 a_s,  c2 = 1/(1 - n.bbar_s),
//         c1 = (n.a_s) c2^2,  bbar_s = (b_s + b_{s+1})/2,  a_s = (b_{s+1} - b_s)/dt
//   near: F_p(g) = sum_s [n/R^2 + i omega (b_s - n)/R] exp(i omega (t_s + R_s))
//                                                                      kernel_nearfield.cl:58-92
//   a step contributes only if |phase_s - phase_{s-1}| < pi (phase_{-1} = 0),
//   kernel_farfield.cl:69-72; snapshots add w dt^2 |F|^2 (or sqrt(w) dt F for
//   the coherent outputs) when it_start + it + 2 hits itSnaps[iSnap],
//   kernel_farfield.cl:100-106.
//
// How it is laid out on gfx950:
//   * A block owns one direction (far) or screen point (near) and a run of
//     T*M consecutive omegas: each of the T = 64*waves threads holds M omegas
//     (k-th omega of thread t is chunk*T*M + k*T + t, so every output store of
//     a wave is 64 consecutive doubles).  A tile index also carries a particle
//     slice q; slices of one tile are combined by a fixed-order reduction.
//   * The block walks its particles (CSR) and their steps in tiles of T steps.
//     Stage: thread t computes, in FP64, everything of step t0+t that does not
//     depend on omega -- tau = t - n.x (far) or t + |P - x| (near) and the
//     amplitude vector(s) -- and stores it in LDS.  Sweep: every thread reads
//     the tile back (broadcast LDS reads) and for each of its omegas forms the
//     phase omega*tau in FP64, applies the |dphase| < pi filter and
//     accumulates.  The omega-independent work is therefore paid once per
//     (point, step) and shared by the block's T*M omegas.
//   * SR_PREC_F64: FP64 everywhere; the phase is reduced modulo 2 pi with a
//     two-constant Cody-Waite step before the FP64 sincos, so the library
//     routine only ever sees |r| <= pi.
//   * SR_PREC_SPLIT: the phase and its reduction stay FP64, the reduced phase
//     (in revolutions, |f| <= 1/2) feeds the hardware v_sin_f32/v_cos_f32,
//     amplitudes are FP32 and the step sums are FP32 with a Kahan fold at the
//     end of every swept segment (so the compensation costs O(1/T) per eval).
//   * No atomics: every output element is owned by exactly one block (or one
//     block per slice, reduced in slice order), so results are deterministic.
//
// All code here is compiled with -ffp-contract=off: the FP64 stage reproduces
// the reference's expression order term by term, and fused multiply-adds are
// written out explicitly where they are wanted (the sweep accumulations).
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "synchrad_amd.h"

namespace {

constexpr int kWave = 64;
constexpr int kMaxT = 256;        // threads per block = steps per LDS tile
constexpr int kMaxSlices = 16;
constexpr int kMaxOut = 6;

constexpr double kPi = 3.141592653589793;
constexpr double kInv2Pi = 0.15915494309189535;
constexpr double k2PiHi = 6.283185307179586;
constexpr double k2PiLo = 2.4492935982947064e-16;

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

struct KParams {
  // tracks (CSR); offsets == nullptr selects the single-particle form
  const int64_t *offsets;
  const double *x, *y, *z, *ux, *uy, *uz;
  const double *weight;
  const uint32_t *it_start, *it_end, *it_snaps;
  uint32_t n_snaps, snaps_stride;
  double dt;
  int64_t single_n;
  double single_w;
  uint32_t single_its, single_ite;
  // grid
  int32_t comp;
  uint32_t n_omega, n_theta, n_phi;
  const double *omega, *sin_t, *cos_t, *radius, *sin_p, *cos_p;
  double l_screen;
  const double *form_factor;
  int64_t n_grid;
  // outputs
  int32_t n_out;
  double *out[kMaxOut];
  double *scratch;  // slices > 1: [slices][n_out][n_grid]
  // tiling
  uint32_t n_chunks;
  uint32_t n_dir;
  int32_t slices;
  int64_t n_tiles;
  int64_t slice_lo[kMaxSlices + 1];
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_reduced(double ph, double *s, double *c) {
  const double k = rint(ph * kInv2Pi);
  double r = fma(-k, k2PiHi, ph);
  r = fma(-k, k2PiLo, r);
  sincos(r, s, c);
}

// 1/sqrt(1 + u.u) as the reference's rsqrt(1 + dot(u,u)), kernel_farfield.cl:77
__device__ __forceinline__ double gamma_inv(double ux, double uy, double uz) {
  return 1.0 / sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
}

struct Particle {
  int64_t off, n;
  double w;
  uint32_t its, ite;
  const uint32_t *snaps;
};

__device__ __forceinline__ Particle load_particle(const KParams &p, int64_t i) {
  Particle r;
  if (p.offsets) {
    r.off = p.offsets[i];
    r.n = p.offsets[i + 1] - r.off;
    r.w = p.weight[i];
    r.its = p.it_start[i];
    r.ite = p.it_end[i];
  } else {
    r.off = 0;
    r.n = p.single_n;
    r.w = p.single_w;
    r.its = p.single_its;
    r.ite = p.single_ite;
  }
  r.snaps = p.it_snaps + i * (int64_t)p.snaps_stride;
  return r;
}

// Snapshot schedule of one particle (kernel_farfield.cl:53-57, 100-106): the
// record for slot iSnap happens after iteration r = itSnaps[iSnap] - 2 - itStart,
// provided r lies in [0, itEnd-2] and after the previous record; a slot that
// can never match stops all later records, as the equality test in the
// reference never fires again.  Reading past the last slot means "no record".
struct SnapCursor {
  uint32_t slot;
  int64_t next;  // -1: no further record
  __device__ void advance(const Particle &pt, uint32_t n_snaps, int64_t last) {
    next = -1;
    if (slot >= n_snaps) return;
    const int64_t r = (int64_t)pt.snaps[slot] - 2 - (int64_t)pt.its;
    if (r <= last || r > (int64_t)pt.ite - 2) return;
    next = r;
  }
  __device__ void init(const Particle &pt, uint32_t n_snaps) {
    slot = 0;
    while (slot < n_snaps && !(pt.its < pt.snaps[slot])) slot++;
    advance(pt, n_snaps, -1);
  }
};

// ---------------------------------------------------------------------------
// the tile kernel
// ---------------------------------------------------------------------------
template <int MODE, int PREC, int M>
__global__ void __launch_bounds__(kMaxT) k_spectrum(const KParams p) {
  using Amp = typename std::conditional<PREC == SR_PREC_SPLIT, float, double>::type;
  constexpr int NA = (MODE == SR_MODE_FAR) ? 3 : 6;   // amplitude words per step
  constexpr int NC = 2 * NA;                           // real accumulators per omega

  __shared__ double s_tau[kMaxT];
  __shared__ Amp s_amp[NA][kMaxT];

  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const double dt = p.dt;
  const double dtInv = 1.0 / dt;
  const bool spheric = (p.comp == SR_COMP_SPHERIC || p.comp == SR_COMP_SPHERIC_COMPLEX);

  for (int64_t tile = blockIdx.x; tile < p.n_tiles; tile += gridDim.x) {
    const uint32_t chunk = (uint32_t)(tile % p.n_chunks);
    const int64_t rest = tile / p.n_chunks;
    const uint32_t dir = (uint32_t)(rest % p.n_dir);
    const int q = (int)(rest / p.n_dir);
    const uint32_t iPhi = dir / p.n_theta;
    const uint32_t iTh = dir - iPhi * p.n_theta;

    // geometry of the block's point (kernel_farfield.cl:40-42, 384-387;
    // kernel_nearfield.cl:40-42)
    double gx, gy, gz;          // far: n ; near: screen point P
    double thx = 0, thy = 0, thz = 0, phx = 0, phy = 0;
    if (MODE == SR_MODE_FAR) {
      const double st = p.sin_t[iTh], ct = p.cos_t[iTh];
      const double sp = p.sin_p[iPhi], cp = p.cos_p[iPhi];
      gx = st * cp;
      gy = st * sp;
      gz = ct;
      thx = ct * cp; thy = ct * sp; thz = -st;
      phx = -sp; phy = cp;
    } else {
      const double r = p.radius[iTh];
      gx = r * p.cos_p[iPhi];
      gy = r * p.sin_p[iPhi];
      gz = p.l_screen;
    }

    int64_t wi[M];
    bool valid[M];
    double om[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      wi[k] = (int64_t)chunk * T * M + (int64_t)k * T + tid;
      valid[k] = wi[k] < (int64_t)p.n_omega;
      om[k] = valid[k] ? p.omega[wi[k]] : 0.0;
    }
    double acc[kMaxOut][M];
#pragma unroll
    for (int c = 0; c < kMaxOut; c++)
#pragma unroll
      for (int k = 0; k < M; k++) acc[c][k] = 0.0;

    const int64_t p_lo = p.slice_lo[q], p_hi = p.slice_lo[q + 1];
    for (int64_t ip = p_lo; ip < p_hi; ip++) {
      const Particle pt = load_particle(p, ip);
      if (pt.n <= 0) continue;
      const int64_t n_act = min(pt.n - 1, (int64_t)pt.ite - 1);
      const double wpdt2 = pt.w * dt * dt;
      const double wpdt = sqrt(pt.w) * dt;

      SnapCursor cur;
      cur.init(pt, p.n_snaps);

      double prev[M];
      // F64: running sums; SPLIT: tile partial P, Kahan sum S and compensation C
      double fs[M][NC];
      float ps[M][NC], ss[M][NC], cs_[M][NC];
#pragma unroll
      for (int k = 0; k < M; k++) {
        prev[k] = 0.0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
          fs[k][c] = 0.0;
          ps[k][c] = 0.f; ss[k][c] = 0.f; cs_[k][c] = 0.f;
        }
      }

      // adds this particle's snapshot value into the block's outputs
      auto emit = [&](uint32_t slot) {
#pragma unroll
        for (int k = 0; k < M; k++) {
          double F[NC];
#pragma unroll
          for (int c = 0; c < NC; c++)
            F[c] = (PREC == SR_PREC_SPLIT) ? ((double)ss[k][c] - (double)cs_[k][c]) : fs[k][c];
          double re[3], im[3];
          if (MODE == SR_MODE_FAR) {
#pragma unroll
            for (int c = 0; c < 3; c++) { re[c] = F[c]; im[c] = F[3 + c]; }
          } else {
            // F = G2 + i omega G1 ; words: G1re 0..2, G1im 3..5, G2re 6..8, G2im 9..11
#pragma unroll
            for (int c = 0; c < 3; c++) {
              re[c] = F[6 + c] - om[k] * F[3 + c];
              im[c] = F[9 + c] + om[k] * F[c];
            }
          }
          double v[kMaxOut];
          switch (p.comp) {
            case SR_COMP_TOTAL:
              v[0] = wpdt2 * ((re[0] * re[0] + re[1] * re[1] + re[2] * re[2]) +
                              (im[0] * im[0] + im[1] * im[1] + im[2] * im[2]));
              break;
            case SR_COMP_CARTESIAN:
            case SR_COMP_SPHERIC:
#pragma unroll
              for (int c = 0; c < 3; c++) v[c] = wpdt2 * (re[c] * re[c] + im[c] * im[c]);
              break;
            default: {
              const double ff = (MODE == SR_MODE_FAR && p.comp == SR_COMP_CARTESIAN_COMPLEX && valid[k])
                                    ? p.form_factor[wi[k]] : 1.0;
#pragma unroll
              for (int c = 0; c < 3; c++) {
                v[2 * c] = wpdt * (ff * re[c]);
                v[2 * c + 1] = wpdt * (ff * im[c]);
              }
            }
          }
          if (p.n_snaps == 1) {
#pragma unroll
            for (int c = 0; c < kMaxOut; c++)
              if (c < p.n_out) acc[c][k] += v[c];
          } else if (valid[k]) {
            const int64_t g = (int64_t)slot * p.n_grid + (int64_t)dir * p.n_omega + wi[k];
#pragma unroll
            for (int c = 0; c < kMaxOut; c++)
              if (c < p.n_out) p.out[c][g] += v[c];
          }
        }
      };

      for (int64_t t0 = 0; t0 < n_act; t0 += T) {
        const int cnt = (int)min((int64_t)T, n_act - t0);
        __syncthreads();  // previous tile fully consumed
        if (tid < cnt) {
          // ---- stage: omega-independent part of step s = t0 + tid, in FP64 ----
          const int64_t s = t0 + tid;
          const int64_t gs = pt.off + s;
          const double time = (double)(uint32_t)(pt.its + (uint32_t)s) * dt;
          const double xs = p.x[gs], ys = p.y[gs], zs = p.z[gs];
          const double u0x = p.ux[gs], u0y = p.uy[gs], u0z = p.uz[gs];
          const double g0 = gamma_inv(u0x, u0y, u0z);
          const double b0x = u0x * g0, b0y = u0y * g0, b0z = u0z * g0;
          if (MODE == SR_MODE_FAR) {
            // kernel_farfield.cl:65-94
            s_tau[tid] = time - (xs * gx + ys * gy + zs * gz);
            const double u1x = p.ux[gs + 1], u1y = p.uy[gs + 1], u1z = p.uz[gs + 1];
            const double g1 = gamma_inv(u1x, u1y, u1z);
            const double b1x = u1x * g1, b1y = u1y * g1, b1z = u1z * g1;
            const double ax = (b1x - b0x) * dtInv, ay = (b1y - b0y) * dtInv, az = (b1z - b0z) * dtInv;
            const double mx = 0.5 * (b1x + b0x), my = 0.5 * (b1y + b0y), mz = 0.5 * (b1z + b0z);
            double c1 = ax * gx + ay * gy + az * gz;
            double c2 = 1.0 - (mx * gx + my * gy + mz * gz);
            c2 = 1.0 / c2;
            c1 = c1 * c2 * c2;
            double Ax = c1 * (gx - mx) - c2 * ax;
            double Ay = c1 * (gy - my) - c2 * ay;
            double Az = c1 * (gz - mz) - c2 * az;
            if (spheric) {
              // kernel_farfield.cl:442-445
              const double r0 = gx * Ax + gy * Ay + gz * Az;
              const double r1 = thx * Ax + thy * Ay + thz * Az;
              const double r2 = phx * Ax + phy * Ay + 0.0 * Az;
              Ax = r0; Ay = r1; Az = r2;
            }
            s_amp[0][tid] = (Amp)Ax;
            s_amp[1][tid] = (Amp)Ay;
            s_amp[2][tid] = (Amp)Az;
          } else {
            // kernel_nearfield.cl:67-88; omega of c1 is applied at the record
            const double rx = gx - xs, ry = gy - ys, rz = gz - zs;
            const double R = sqrt(rx * rx + ry * ry + rz * rz);
            s_tau[tid] = time + R;
            const double rInv = 1.0 / R;
            const double nx = rInv * rx, ny = rInv * ry, nz = rInv * rz;
            const double r2 = rInv * rInv;
            s_amp[0][tid] = (Amp)(rInv * (b0x - nx));
            s_amp[1][tid] = (Amp)(rInv * (b0y - ny));
            s_amp[2][tid] = (Amp)(rInv * (b0z - nz));
            s_amp[3][tid] = (Amp)(r2 * nx);
            s_amp[4][tid] = (Amp)(r2 * ny);
            s_amp[5][tid] = (Amp)(r2 * nz);
          }
        }
        __syncthreads();

        // ---- sweep, split at snapshot records falling inside the tile ----
        int j = 0;
        while (j < cnt) {
          const bool rec_here = (cur.next >= 0) && (cur.next < t0 + cnt);
          // records are strictly increasing, so cur.next >= t0 + j; the max()
          // only guarantees forward progress
          const int jend = rec_here ? max(j + 1, (int)(cur.next - t0) + 1) : cnt;
          for (; j < jend; j++) {
            const double tau = s_tau[j];
            Amp a[NA];
#pragma unroll
            for (int c = 0; c < NA; c++) a[c] = s_amp[c][j];
#pragma unroll
            for (int k = 0; k < M; k++) {
              const double ph = om[k] * tau;
              const double d = ph - prev[k];
              prev[k] = ph;
              const bool act = valid[k] && (fabs(d) < kPi);
              if (act) {
                if (PREC == SR_PREC_SPLIT) {
                  const double rv = ph * kInv2Pi;
                  const float f = (float)(rv - rint(rv));
                  const float sn = __builtin_amdgcn_sinf(f);
                  const float cn = __builtin_amdgcn_cosf(f);
                  // word layout shared with the FP64 path: per amplitude vector
                  // v (far: A; near: v1 then v2) [v cos | v sin], 3 words each
#pragma unroll
                  for (int c = 0; c < NA; c++) {
                    const int w = 6 * (c / 3) + (c % 3);
                    ps[k][w] = __builtin_fmaf((float)a[c], cn, ps[k][w]);
                    ps[k][w + 3] = __builtin_fmaf((float)a[c], sn, ps[k][w + 3]);
                  }
                } else {
                  double sn, cn;
                  sincos_reduced(ph, &sn, &cn);
                  if (MODE == SR_MODE_FAR) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                      fs[k][c] = fma((double)a[c], cn, fs[k][c]);
                      fs[k][3 + c] = fma((double)a[c], sn, fs[k][3 + c]);
                    }
                  } else {
                    // G1 += v1 e^{i ph}, G2 += v2 e^{i ph}
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                      fs[k][c] = fma((double)a[c], cn, fs[k][c]);
                      fs[k][3 + c] = fma((double)a[c], sn, fs[k][3 + c]);
                      fs[k][6 + c] = fma((double)a[3 + c], cn, fs[k][6 + c]);
                      fs[k][9 + c] = fma((double)a[3 + c], sn, fs[k][9 + c]);
                    }
                  }
                }
              }
            }
          }
          if (PREC == SR_PREC_SPLIT) {
            // Kahan fold of the segment's partial sums
#pragma unroll
            for (int k = 0; k < M; k++)
#pragma unroll
              for (int c = 0; c < NC; c++) {
                const float yv = ps[k][c] - cs_[k][c];
                const float tv = ss[k][c] + yv;
                cs_[k][c] = (tv - ss[k][c]) - yv;
                ss[k][c] = tv;
                ps[k][c] = 0.f;
              }
          }
          if (rec_here) {
            emit(cur.slot);
            const int64_t last = cur.next;
            cur.slot++;
            cur.advance(pt, p.n_snaps, last);
          }
        }
      }
      // records after the last active step (global it_range mode) see the final sums
      while (cur.next >= 0) {
        emit(cur.slot);
        const int64_t last = cur.next;
        cur.slot++;
        cur.advance(pt, p.n_snaps, last);
      }
    }

    if (p.n_snaps == 1) {
#pragma unroll
      for (int k = 0; k < M; k++) {
        if (!valid[k]) continue;
        const int64_t g = (int64_t)dir * p.n_omega + wi[k];
#pragma unroll
        for (int c = 0; c < kMaxOut; c++) {
          if (c >= p.n_out) break;
          if (p.slices > 1)
            p.scratch[((int64_t)q * p.n_out + c) * p.n_grid + g] = acc[c][k];
          else
            p.out[c][g] += acc[c][k];
        }
      }
    }
  }
}

// out[c][i] += sum_q scratch[q][c][i], q in order
__global__ void k_reduce_slices(const KParams p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)p.n_out * p.n_grid;
  if (i >= total) return;
  const int c = (int)(i / p.n_grid);
  const int64_t g = i - (int64_t)c * p.n_grid;
  double s = 0.0;
  for (int q = 0; q < p.slices; q++) s += p.scratch[((int64_t)q * p.n_out + c) * p.n_grid + g];
  p.out[c][g] += s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int outputs_for(int32_t mode, int32_t comp) {
  if (mode == SR_MODE_FAR) {
    switch (comp) {
      case SR_COMP_TOTAL: return 1;
      case SR_COMP_CARTESIAN: case SR_COMP_SPHERIC: return 3;
      case SR_COMP_CARTESIAN_COMPLEX: case SR_COMP_SPHERIC_COMPLEX: return 6;
    }
  } else if (mode == SR_MODE_NEAR) {
    switch (comp) {
      case SR_COMP_TOTAL: return 1;
      case SR_COMP_CARTESIAN: return 3;
      case SR_COMP_CARTESIAN_COMPLEX: return 6;
    }
  }
  return -1;
}

struct Plan {
  int waves;     // waves per block
  int M;         // omegas per thread
  int slices;
  int64_t blocks;
};

// Tile shape: prefer the largest block whose omega run is mostly used,
// then enough particle slices to put ~8 waves on each of the 1024 SIMDs.
Plan make_plan(const sr_grid &g, const sr_opts &o, int64_t n_particles, uint32_t n_snaps) {
  Plan pl;
  const int64_t n_om = g.n_omega;
  const int64_t n_dir = (int64_t)g.n_theta * g.n_phi;
  int best_w = 1, best_m = 1;
  double best_score = -1;
  const int m_opts_split[2] = {1, 2};
  for (int w = 1; w <= 4; w *= 2) {
    for (int mi = 0; mi < 2; mi++) {
      const int m = m_opts_split[mi];
      if (o.precision == SR_PREC_F64 && m > 1) continue;
      const int64_t run = (int64_t)kWave * w * m;
      const int64_t chunks = (n_om + run - 1) / run;
      const double eff = (double)n_om / (double)(chunks * run);
      if (eff < 0.75) continue;
      const double score = eff + 0.01 * w + 0.005 * m;
      if (score > best_score) { best_score = score; best_w = w; best_m = m; }
    }
  }
  if (best_score < 0) { best_w = 1; best_m = 1; }
  pl.waves = o.waves_per_block > 0 ? o.waves_per_block : best_w;
  pl.M = o.omega_per_lane > 0 ? o.omega_per_lane : best_m;
  if (o.precision == SR_PREC_F64) pl.M = 1;
  const int64_t run = (int64_t)kWave * pl.waves * pl.M;
  const int64_t chunks = (n_om + run - 1) / run;
  const int64_t base_waves = chunks * n_dir * pl.waves;
  int s = 1;
  if (o.slices > 0) {
    s = o.slices;
  } else if (n_snaps == 1) {
    const int64_t target = 8 * 1024;
    while (s < kMaxSlices && base_waves * s < target && (int64_t)(s * 2) <= n_particles) s *= 2;
  }
  if (n_snaps != 1) s = 1;
  if (s > n_particles) s = (int)std::max<int64_t>(1, n_particles);
  pl.slices = s;
  pl.blocks = chunks * n_dir * s;
  return pl;
}

template <int MODE, int PREC, int M>
hipError_t launch_t(const KParams &kp, int64_t grid, int threads, hipStream_t st) {
  hipLaunchKernelGGL((k_spectrum<MODE, PREC, M>), dim3((unsigned)grid), dim3(threads), 0, st, kp);
  return hipGetLastError();
}

hipError_t launch_tiles(int mode, int prec, int M, const KParams &kp, int64_t grid, int threads,
                        hipStream_t st) {
  if (mode == SR_MODE_FAR) {
    if (prec == SR_PREC_F64) return launch_t<SR_MODE_FAR, SR_PREC_F64, 1>(kp, grid, threads, st);
    if (M == 2) return launch_t<SR_MODE_FAR, SR_PREC_SPLIT, 2>(kp, grid, threads, st);
    return launch_t<SR_MODE_FAR, SR_PREC_SPLIT, 1>(kp, grid, threads, st);
  }
  if (prec == SR_PREC_F64) return launch_t<SR_MODE_NEAR, SR_PREC_F64, 1>(kp, grid, threads, st);
  if (M == 2) return launch_t<SR_MODE_NEAR, SR_PREC_SPLIT, 2>(kp, grid, threads, st);
  return launch_t<SR_MODE_NEAR, SR_PREC_SPLIT, 1>(kp, grid, threads, st);
}

int check_grid(const sr_grid *g) {
  if (!g) return fail(SR_ERR_INVALID, "grid is NULL");
  if (g->mode != SR_MODE_FAR && g->mode != SR_MODE_NEAR)
    return fail(SR_ERR_INVALID, "grid.mode %d is not far (0) or near (1)", g->mode);
  if (g->n_omega == 0 || g->n_theta == 0 || g->n_phi == 0)
    return fail(SR_ERR_INVALID, "grid sizes must be >= 1 (got %u, %u, %u)", g->n_omega, g->n_theta,
                g->n_phi);
  if (!g->omega || !g->sin_phi || !g->cos_phi) return fail(SR_ERR_INVALID, "grid axis pointer is NULL");
  if (g->mode == SR_MODE_FAR && (!g->sin_theta || !g->cos_theta))
    return fail(SR_ERR_INVALID, "far grid needs sin_theta and cos_theta");
  if (g->mode == SR_MODE_NEAR && !g->radius) return fail(SR_ERR_INVALID, "near grid needs radius");
  const double npts = (double)g->n_omega * g->n_theta * g->n_phi;
  if (npts > 4.0e9) return fail(SR_ERR_INVALID, "grid has more than 4e9 points");
  return SR_OK;
}

int check_opts(const sr_grid *g, const sr_opts *o) {
  if (!o) return fail(SR_ERR_INVALID, "opts is NULL");
  if (outputs_for(g->mode, o->comp) < 0)
    return fail(SR_ERR_INVALID, "comp %d is not available in mode %d", o->comp, g->mode);
  if (o->precision != SR_PREC_F64 && o->precision != SR_PREC_SPLIT)
    return fail(SR_ERR_INVALID, "precision %d is not F64 (0) or SPLIT (1)", o->precision);
  if (o->omega_per_lane < 0 || o->omega_per_lane > 2)
    return fail(SR_ERR_INVALID, "omega_per_lane must be 0, 1 or 2");
  if (!(o->waves_per_block == 0 || o->waves_per_block == 1 || o->waves_per_block == 2 ||
        o->waves_per_block == 4))
    return fail(SR_ERR_INVALID, "waves_per_block must be 0, 1, 2 or 4");
  if (o->slices < 0 || o->slices > kMaxSlices)
    return fail(SR_ERR_INVALID, "slices must be in [0, %d]", kMaxSlices);
  if (o->persistent < 0) return fail(SR_ERR_INVALID, "persistent must be >= 0");
  if (g->mode == SR_MODE_FAR && o->comp == SR_COMP_CARTESIAN_COMPLEX && !o->form_factor)
    return fail(SR_ERR_INVALID, "far cartesian_complex needs form_factor");
  return SR_OK;
}

void fill_grid(KParams &kp, const sr_grid &g, const sr_opts &o, int n_out, double *const *outs) {
  kp.comp = o.comp;
  kp.n_omega = g.n_omega;
  kp.n_theta = g.n_theta;
  kp.n_phi = g.n_phi;
  kp.omega = g.omega;
  kp.sin_t = g.sin_theta;
  kp.cos_t = g.cos_theta;
  kp.radius = g.radius;
  kp.sin_p = g.sin_phi;
  kp.cos_p = g.cos_phi;
  kp.l_screen = g.l_screen;
  kp.form_factor = o.form_factor;
  kp.n_grid = (int64_t)g.n_omega * g.n_theta * g.n_phi;
  kp.n_out = n_out;
  for (int c = 0; c < kMaxOut; c++) kp.out[c] = c < n_out ? outs[c] : nullptr;
  kp.n_dir = g.n_theta * g.n_phi;
}

// balanced contiguous particle slices by step count
void slice_particles(KParams &kp, int slices, int64_t n_particles, const int64_t *offsets_host) {
  kp.slice_lo[0] = 0;
  if (offsets_host && slices > 1) {
    const int64_t total = offsets_host[n_particles] - offsets_host[0];
    int64_t i = 0;
    for (int q = 1; q < slices; q++) {
      const double target = (double)total * q / slices + (double)offsets_host[0];
      while (i < n_particles && (double)offsets_host[i] < target) i++;
      kp.slice_lo[q] = std::max(kp.slice_lo[q - 1], i);
    }
  } else {
    for (int q = 1; q < slices; q++) kp.slice_lo[q] = n_particles * q / slices;
  }
  kp.slice_lo[slices] = n_particles;
}

int run(KParams &kp, const Plan &pl, const sr_grid &g, const sr_opts &o, void *ws, size_t ws_bytes,
        hipStream_t st) {
  const int threads = kWave * pl.waves;
  const int64_t run = (int64_t)threads * pl.M;
  kp.n_chunks = (uint32_t)((g.n_omega + run - 1) / run);
  kp.slices = pl.slices;
  kp.n_tiles = pl.blocks;
  if (pl.slices > 1) {
    const size_t need = (size_t)pl.slices * kp.n_out * kp.n_grid * sizeof(double);
    if (!ws || ws_bytes < need)
      return fail(SR_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, ws_bytes);
    kp.scratch = (double *)ws;
  } else {
    kp.scratch = nullptr;
  }
  int64_t grid = pl.blocks;
  if (o.persistent > 0) grid = std::min<int64_t>(grid, o.persistent);
  if (grid <= 0) return SR_OK;
  hipError_t e = launch_tiles(g.mode, o.precision, pl.M, kp, grid, threads, st);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "tile kernel launch: %s", hipGetErrorString(e));
  if (pl.slices > 1) {
    const int64_t total = (int64_t)kp.n_out * kp.n_grid;
    const int tb = 256;
    hipLaunchKernelGGL(k_reduce_slices, dim3((unsigned)((total + tb - 1) / tb)), dim3(tb), 0, st, kp);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SR_ERR_HIP, "slice reduction launch: %s", hipGetErrorString(e));
  }
  return SR_OK;
}

int check_tracks(const sr_tracks *t) {
  if (!t) return fail(SR_ERR_INVALID, "tracks is NULL");
  if (t->n_particles < 0 || t->n_points < 0) return fail(SR_ERR_INVALID, "negative track sizes");
  if (t->n_particles == 0) return SR_OK;
  if (!t->offsets || !t->x || !t->y || !t->z || !t->ux || !t->uy || !t->uz || !t->weight ||
      !t->it_start || !t->it_end || !t->it_snaps)
    return fail(SR_ERR_INVALID, "a track pointer is NULL");
  if (t->n_snaps == 0) return fail(SR_ERR_INVALID, "n_snaps must be >= 1");
  if (t->snaps_stride != 0 && t->snaps_stride != t->n_snaps)
    return fail(SR_ERR_INVALID, "snaps_stride must be 0 or n_snaps");
  if (!(t->dt != 0.0) || !isfinite(t->dt)) return fail(SR_ERR_INVALID, "dt must be finite and non-zero");
  return SR_OK;
}

int single_particle(int32_t mode, int32_t comp, double *const *outs, const double *x, const double *y,
                    const double *z, const double *ux, const double *uy, const double *uz, double wp,
                    uint32_t itStart, uint32_t itEnd, uint32_t nSteps, const sr_grid &g, double dt,
                    uint32_t nSnaps, const uint32_t *itSnaps, int32_t precision, void *stream) {
  int rc = check_grid(&g);
  if (rc) return rc;
  sr_opts o;
  memset(&o, 0, sizeof(o));
  o.comp = comp;
  o.precision = precision;
  if ((rc = check_opts(&g, &o))) return rc;
  // the reference would loop 2^32 times (itEnd == 0) or read out of bounds (nSteps == 0)
  if (itEnd == 0) return fail(SR_ERR_INVALID, "itEnd must be >= 1");
  if (nSteps == 0) return fail(SR_ERR_INVALID, "nSteps must be >= 1");
  if (nSnaps == 0 || !itSnaps) return fail(SR_ERR_INVALID, "nSnaps must be >= 1 with itSnaps");
  if (!x || !y || !z || !ux || !uy || !uz) return fail(SR_ERR_INVALID, "a track pointer is NULL");
  if (!(dt != 0.0) || !isfinite(dt)) return fail(SR_ERR_INVALID, "dt must be finite and non-zero");
  const int n_out = outputs_for(mode, comp);
  for (int c = 0; c < n_out; c++)
    if (!outs[c]) return fail(SR_ERR_INVALID, "output %d is NULL", c);
  KParams kp;
  memset(&kp, 0, sizeof(kp));
  kp.offsets = nullptr;
  kp.x = x; kp.y = y; kp.z = z; kp.ux = ux; kp.uy = uy; kp.uz = uz;
  kp.it_snaps = itSnaps;
  kp.n_snaps = nSnaps;
  kp.snaps_stride = 0;
  kp.dt = dt;
  kp.single_n = nSteps;
  kp.single_w = wp;
  kp.single_its = itStart;
  kp.single_ite = itEnd;
  fill_grid(kp, g, o, n_out, outs);
  Plan pl = make_plan(g, o, 1, nSnaps);
  pl.slices = 1;  // one particle: no slices
  slice_particles(kp, 1, 1, nullptr);
  return run(kp, pl, g, o, nullptr, 0, (hipStream_t)stream);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int sr_abi_version(void) { return SR_ABI_VERSION; }

const char *sr_last_error(void) { return g_err; }

int sr_num_outputs(int32_t mode, int32_t comp) {
  const int n = outputs_for(mode, comp);
  return n < 0 ? fail(SR_ERR_INVALID, "comp %d is not available in mode %d", comp, mode) : n;
}

int sr_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                       size_t *bytes) {
  int rc;
  if (!bytes) return fail(SR_ERR_INVALID, "bytes is NULL");
  *bytes = 0;
  if ((rc = check_tracks(tracks)) || (rc = check_grid(grid)) || (rc = check_opts(grid, opts))) return rc;
  const Plan pl = make_plan(*grid, *opts, tracks->n_particles, tracks->n_snaps);
  if (pl.slices > 1)
    *bytes = (size_t)pl.slices * outputs_for(grid->mode, opts->comp) *
             ((size_t)grid->n_omega * grid->n_theta * grid->n_phi) * sizeof(double);
  return SR_OK;
}

int sr_spectrum_batch(const sr_tracks *t, const sr_grid *g, const sr_opts *o, double *const *outs,
                      void *workspace, size_t workspace_bytes, void *stream) {
  int rc;
  if ((rc = check_tracks(t)) || (rc = check_grid(g)) || (rc = check_opts(g, o))) return rc;
  if (t->n_particles == 0) return SR_OK;
  const int n_out = outputs_for(g->mode, o->comp);
  if (!outs) return fail(SR_ERR_INVALID, "outs is NULL");
  for (int c = 0; c < n_out; c++)
    if (!outs[c]) return fail(SR_ERR_INVALID, "output %d is NULL", c);
  KParams kp;
  memset(&kp, 0, sizeof(kp));
  kp.offsets = t->offsets;
  kp.x = t->x; kp.y = t->y; kp.z = t->z; kp.ux = t->ux; kp.uy = t->uy; kp.uz = t->uz;
  kp.weight = t->weight;
  kp.it_start = t->it_start;
  kp.it_end = t->it_end;
  kp.it_snaps = t->it_snaps;
  kp.n_snaps = t->n_snaps;
  kp.snaps_stride = t->snaps_stride;
  kp.dt = t->dt;
  fill_grid(kp, *g, *o, n_out, outs);
  const Plan pl = make_plan(*g, *o, t->n_particles, t->n_snaps);
  slice_particles(kp, pl.slices, t->n_particles, t->offsets_host);
  return run(kp, pl, *g, *o, workspace, workspace_bytes, (hipStream_t)stream);
}

int sr_far_total(double *spectrum, const double *x, const double *y, const double *z,
                 const double *ux, const double *uy, const double *uz, double wp, uint32_t itStart,
                 uint32_t itEnd, uint32_t nSteps, const double *omega, const double *sinTheta,
                 const double *cosTheta, const double *sinPhi, const double *cosPhi, uint32_t nOmega,
                 uint32_t nTheta, uint32_t nPhi, double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                 int32_t precision, void *stream) {
  sr_grid g;
  memset(&g, 0, sizeof(g));
  g.mode = SR_MODE_FAR;
  g.n_omega = nOmega; g.n_theta = nTheta; g.n_phi = nPhi;
  g.omega = omega; g.sin_theta = sinTheta; g.cos_theta = cosTheta; g.sin_phi = sinPhi; g.cos_phi = cosPhi;
  double *outs[1] = {spectrum};
  return single_particle(SR_MODE_FAR, SR_COMP_TOTAL, outs, x, y, z, ux, uy, uz, wp, itStart, itEnd,
                         nSteps, g, dt, nSnaps, itSnaps, precision, stream);
}

int sr_far_cartesian(double *spectrum1, double *spectrum2, double *spectrum3, const double *x,
                     const double *y, const double *z, const double *ux, const double *uy,
                     const double *uz, double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                     const double *omega, const double *sinTheta, const double *cosTheta,
                     const double *sinPhi, const double *cosPhi, uint32_t nOmega, uint32_t nTheta,
                     uint32_t nPhi, double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                     int32_t precision, void *stream) {
  sr_grid g;
  memset(&g, 0, sizeof(g));
  g.mode = SR_MODE_FAR;
  g.n_omega = nOmega; g.n_theta = nTheta; g.n_phi = nPhi;
  g.omega = omega; g.sin_theta = sinTheta; g.cos_theta = cosTheta; g.sin_phi = sinPhi; g.cos_phi = cosPhi;
  double *outs[3] = {spectrum1, spectrum2, spectrum3};
  return single_particle(SR_MODE_FAR, SR_COMP_CARTESIAN, outs, x, y, z, ux, uy, uz, wp, itStart, itEnd,
                         nSteps, g, dt, nSnaps, itSnaps, precision, stream);
}

int sr_near_total(double *spectrum, const double *x, const double *y, const double *z,
                  const double *ux, const double *uy, const double *uz, double wp, uint32_t itStart,
                  uint32_t itEnd, uint32_t nSteps, const double *omega, const double *radius,
                  const double *sinPhi, const double *cosPhi, double distanceToScreen, uint32_t nOmega,
                  uint32_t nRadius, uint32_t nPhi, double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                  int32_t precision, void *stream) {
  sr_grid g;
  memset(&g, 0, sizeof(g));
  g.mode = SR_MODE_NEAR;
  g.n_omega = nOmega; g.n_theta = nRadius; g.n_phi = nPhi;
  g.omega = omega; g.radius = radius; g.sin_phi = sinPhi; g.cos_phi = cosPhi;
  g.l_screen = distanceToScreen;
  double *outs[1] = {spectrum};
  return single_particle(SR_MODE_NEAR, SR_COMP_TOTAL, outs, x, y, z, ux, uy, uz, wp, itStart, itEnd,
                         nSteps, g, dt, nSnaps, itSnaps, precision, stream);
}

int sr_near_cartesian(double *spectrum1, double *spectrum2, double *spectrum3, const double *x,
                      const double *y, const double *z, const double *ux, const double *uy,
                      const double *uz, double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                      const double *omega, const double *radius, const double *sinPhi,
                      const double *cosPhi, double distanceToScreen, uint32_t nOmega, uint32_t nRadius,
                      uint32_t nPhi, double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                      int32_t precision, void *stream) {
  sr_grid g;
  memset(&g, 0, sizeof(g));
  g.mode = SR_MODE_NEAR;
  g.n_omega = nOmega; g.n_theta = nRadius; g.n_phi = nPhi;
  g.omega = omega; g.radius = radius; g.sin_phi = sinPhi; g.cos_phi = cosPhi;
  g.l_screen = distanceToScreen;
  double *outs[3] = {spectrum1, spectrum2, spectrum3};
  return single_particle(SR_MODE_NEAR, SR_COMP_CARTESIAN, outs, x, y, z, ux, uy, uz, wp, itStart, itEnd,
                         nSteps, g, dt, nSnaps, itSnaps, precision, stream);
}

}  // extern "C"
