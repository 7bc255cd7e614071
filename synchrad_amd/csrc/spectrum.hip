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
//     T*M omegas: each of the T = 64*waves threads holds M omegas (the k-th
//     omega of thread t is chunk*T*M + k*T + t', t' = t with the 64-omega runs
//     dealt to the waves in an order rotated by the tile index, so every
//     output access of a wave covers 64 consecutive doubles).  A tile index
//     also carries a particle slice q; the slices of one tile are summed in
//     slice order by a second kernel.
//   * A pre-pass (k_step_prep) computes the direction-independent data of
//     every sample once per call: far a and the mean velocity of the interval,
//     near the velocity.
//   * The block walks its particles (CSR) and their steps in tiles of T steps.
//     Stage: thread t computes, in FP64, everything of step t0+t that does not
//     depend on omega -- tau = t - n.x (far) or t + |P - x| (near) and the
//     amplitude words -- and stores it in LDS.  Sweep: every thread reads
//     the tile back with broadcast LDS reads and accumulates its omegas.  The
//     omega-independent work is paid once per (point, step) and shared by the
//     block's T*M omegas.
//   * Amplitude words: far A.e_theta, A.e_phi (A is perpendicular to n; the
//     record rebuilds x, y, z); near v1 = (b - n)/R and v2 = n/R^2.
//   * Sweeps:
//       F64 direct    FP64 phase omega*tau, |dphase| < pi filter exactly as the
//                     reference forms it, FP64 sin/cos (pi/128 reduction, an
//                     LDS table of sin/cos(i pi/128), short polynomials), FP64
//                     sums; 1 or 2 omegas per thread.
//       split direct  (grids without a constant step) FP64 phase in
//                     revolutions -> v_fract_f64 -> v_sin_f32/v_cos_f32, the
//                     filter in its threshold form |omega|/2pi < 1/(2|dtau|),
//                     FP32 sums with a Kahan fold per swept segment.
//       split rec     (linear grids) the thread's M omegas differ by T*domega,
//                     so exp(i omega_k tau) is advanced from k to k+1 by one
//                     complex multiply with exp(i T domega tau) (computed per
//                     step in the stage).  The tile's steps are put in an order
//                     in which every 64-omega run finds the steps keeping any of
//                     its omegas first (a stable counting sort), so each wave
//                     sweeps only those, stops the rotation chain at its last
//                     kept run and masks only that run.
//   * Records: with one snapshot and one output each thread sums its points
//     over the block's particles in LDS and writes once; otherwise outputs are
//     read-modify-written per particle record by the one thread that owns
//     them.  No atomics anywhere, so results do not depend on scheduling.
//   * An internal output kind writes each particle's complex field separately
//     (the adjoint's forward pass).
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
#include <mutex>
#include <type_traits>

#include "synchrad_amd.h"
#include "common.h"

namespace sr {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

const char *last_error() { return g_err; }

}  // namespace sr

namespace {

using namespace sr;

enum { SWEEP_DIRECT = 0, SWEEP_REC = 1 };

template <int V>
using IC = std::integral_constant<int, V>;

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
  const double *screen_y;   // near, Cartesian screen: P = (radius[i], screen_y[j], L)
  double l_screen;
  double omega_step;        // rec sweep: omega[i+1] - omega[i]
  const double *form_factor;
  int64_t n_grid;
  // outputs (slices > 1: scratch[q][c] replaces out[c])
  int32_t n_out;
  double *out[kMaxOut];
  double *scratch;
  // 1: one snapshot and one output -- each thread sums its points over the
  // block's particles in LDS and adds the sum to the output once
  int32_t acc_lds;
  // per-particle outputs (the adjoint's fields): particle ip writes at
  // (ip - part_base) * part_stride past the usual index; 0 otherwise
  int64_t part_base, part_stride;
  // direction-independent step data from k_step_prep, or nullptr (the stage
  // then computes it), one record per sample: far [n_points][a_x, a_y, a_z,
  // bbar_x, bbar_y, bbar_z] (48 B, 16-byte aligned), near [n_points][b_x, b_y,
  // b_z].  One base address and fixed offsets: the stage reads a record with
  // three 16-byte loads and keeps no per-array base registers
  const double *pre;
  int64_t n_points;
  // tiling
  uint32_t n_chunks;
  uint32_t n_dir;
  UDiv div_chunks, div_dir, div_theta;   // by n_chunks, n_dir, n_theta (tiles < 2^31)
  int32_t slices;
  int64_t n_tiles;
  int64_t slice_lo[kMaxSlices + 1];
  // launch order of the tiles (heaviest first), or nullptr: index order
  const int32_t *tile_order;
};

// ---------------------------------------------------------------------------
// the tile kernel
// ---------------------------------------------------------------------------
// Dynamic LDS of the split sweeps: the Kahan sum S and compensation C of every
// thread's NC*M accumulators, [2][NC*M][T] floats; a thread only ever touches
// its own column, so no barrier guards them.
template <int MODE, int M>
constexpr size_t kahan_lds_bytes(int T) {
  return (size_t)2 * (MODE == SR_MODE_FAR ? 4 : 6) * M * T * sizeof(float);
}

__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return fma(a, b, c); }

// One step's contribution to one frequency's sums acc[NC], from the step's
// amplitude words a[NA] and e^{i phase} = cn + i sn.  Far: [A cos | A sin]
// with NA words each (split: A.e_theta, A.e_phi; FP64: A_x, A_y, A_z).
// Near: the field itself, F_c += (v2_c + i omega v1_c) e^{i phase} with
// a = [v1 (3) | v2 (3)] and w = omega (kernel_nearfield.cl:87-91: c1 =
// omega v1, c2 = v2), as [Re F (3) | Im F (3)].
template <int MODE, int NA, typename X>
__device__ __forceinline__ void accumulate(X *acc, const X *a, X cn, X sn, X w) {
  if constexpr (MODE == SR_MODE_FAR) {
#pragma unroll
    for (int c = 0; c < NA; c++) {
      acc[c] = fmx(a[c], cn, acc[c]);
      acc[NA + c] = fmx(a[c], sn, acc[NA + c]);
    }
  } else {
    const X wc = w * cn, ws = w * sn;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[c] = fmx(a[3 + c], cn, fmx(-a[c], ws, acc[c]));
      acc[3 + c] = fmx(a[3 + c], sn, fmx(a[c], wc, acc[3 + c]));
    }
  }
}
// The split sweeps' register partials of one frequency.  Far: as accumulate.
// Near: the two amplitude parts kept apart, [v2 cos | v2 sin | v1 sin | v1 cos]
// (12 sums, 12 FMAs per step and no multiplication by omega); the Kahan fold
// forms the field's [Re F | Im F] = [v2 cos - omega v1 sin | v2 sin + omega
// v1 cos] from them once per tile.
template <int MODE, int NA>
__device__ __forceinline__ void split_acc(float *acc, const float *a, float cn, float sn) {
  if constexpr (MODE == SR_MODE_FAR) {
#pragma unroll
    for (int c = 0; c < NA; c++) {
      acc[c] = fmx(a[c], cn, acc[c]);
      acc[NA + c] = fmx(a[c], sn, acc[NA + c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      acc[c] = fmx(a[3 + c], cn, acc[c]);
      acc[3 + c] = fmx(a[3 + c], sn, acc[3 + c]);
      acc[6 + c] = fmx(a[c], sn, acc[6 + c]);
      acc[9 + c] = fmx(a[c], cn, acc[9 + c]);
    }
  }
}

// Dynamic LDS after the Kahan state: the per-thread particle sums of the
// single output when KParams::acc_lds is set, [M][T] doubles.
template <int M>
constexpr size_t acc_lds_bytes(int T) {
  return (size_t)M * T * sizeof(double);
}

// amplitude words per staged step: far A.e_theta, A.e_phi; near v1 and v2
// (three components each)
template <int MODE, int PREC>
constexpr int amp_words() {
  return MODE == SR_MODE_FAR ? 2 : 6;
}

// Dynamic LDS before the Kahan state: one tile of T staged steps -- tau
// (double), the split threshold (double), the recurrence rotation (2 floats)
// and the amplitude words -- sized by the launched block, so that one-wave
// blocks do not hold a 256-step tile's LDS.
template <int MODE, int PREC, int SWEEP>
constexpr size_t stage_lds_bytes(int T) {
  constexpr bool split = PREC == SR_PREC_SPLIT;
  return (size_t)T * (sizeof(double) + (split ? sizeof(double) : 0) + (SWEEP == SWEEP_REC ? 2 * sizeof(float) : 0) +
                      amp_words<MODE, PREC>() * (split ? sizeof(float) : sizeof(double)));
}

// one staged step of the recurrence sweep (split): tau, threshold, the
// rotation e^{i T domega tau} and the amplitude words in one 16-byte aligned
// record (32 B far, 48 B near), so that the sweep reads a group of steps from
// one base address with immediate offsets (the same bytes as the separate
// arrays of the other sweeps)
template <int NA>
struct alignas(16) RecStep {
  double tau, thr;
  float er, ei;
  float amp[NA];
};

// rows of a [n][T] LDS array: rows[c][j]
template <typename X>
struct LdsRows {
  X *base;
  int T;
  __device__ X *operator[](int c) const { return base + c * T; }
};

// Register budget per variant, as the minimum number of waves each SIMD must
// be able to hold: the split sweeps keep their Kahan state in LDS, so their
// registers hold the partial sums of M omegas plus one staged step group.
// (far recurrence with M = 2: 4 waves, measured 3% faster than 3 despite a
// few spilled stage registers; the other recurrence variants spill heavily
// at 4 and run at 2)
#ifndef SR_REC_WAVES
#define SR_REC_WAVES 3
#endif
#ifndef SR_REC_WAVES_FAR2
#define SR_REC_WAVES_FAR2 4
#endif
#ifndef SR_REC_WAVES_FAR4
#define SR_REC_WAVES_FAR4 2
#endif
// recurrence groups of 4 steps: all base sin/cos first, then the sums
#ifndef SR_REC_PIPE
#define SR_REC_PIPE 1
#endif
#ifndef SR_WAVE_ROTATE
#define SR_WAVE_ROTATE 1
#endif
#ifndef SR_DIRECT_WAVES
#define SR_DIRECT_WAVES 3
#endif
// split stage: hardware FP64 reciprocal instead of IEEE division
#ifndef SR_SPLIT_RCP
#define SR_SPLIT_RCP 1
#endif
// split direct sweep in 4-step groups
#ifndef SR_DIRECT_GROUP
#define SR_DIRECT_GROUP 1
#endif
#ifndef SR_F64_WAVES
#define SR_F64_WAVES 4
#endif
// steps per iteration of the direct sweep's step loop (FP64 and the split tail)
#ifndef SR_F64_UNROLL
#define SR_F64_UNROLL 4
#endif
#ifndef SR_F64_WAVES2
#define SR_F64_WAVES2 4
#endif
#ifndef SR_F64_NEAR_WAVES2
#define SR_F64_NEAR_WAVES2 SR_F64_WAVES2
#endif
// Recurrence sweep: the steps of a tile are put in an order in which every
// 64-omega run of the block finds the steps that keep any of its omegas first,
// so a wave sweeps only those (a stable counting sort per tile).  The direct
// sweep measured slower with it (the sort's registers cost it a wave per SIMD
// and skipped steps are already cheap there), so it keeps the tile order.
// recurrence sweep: a thread's third and fourth omegas by the three-term
// recurrence of equally spaced angles instead of a complex multiplication
#ifndef SR_REC_CHEB
#define SR_REC_CHEB 1
#endif
#ifndef SR_COMPACT
#define SR_COMPACT 1
#endif
// recurrence sweep: a step's previous retarded time tau_{s-1} comes from the
// neighbouring thread's stage through LDS (1) instead of being recomputed
// from the previous sample (0); the tile's first thread recomputes it.  Same
// results; three global loads fewer per step (the HBM traffic is unchanged:
// the neighbouring lane's load had brought those lines in)
#ifndef SR_TAU_EXCHANGE
#define SR_TAU_EXCHANGE 1
#endif
// 1: a step's bucket (the number of runs it keeps) by binary search over the
// runs' sorted smallest |omega| (log2 R compares; +0.8 % on the betatron
// launch, same buckets); 0: by comparing with every run
#ifndef SR_BUCKET_SEARCH
#define SR_BUCKET_SEARCH 1
#endif

template <int MODE, int PREC, int SWEEP, int M>
struct WavesPerEU {
  static constexpr int value =
      (PREC == SR_PREC_F64) ? (M == 2 ? (MODE == SR_MODE_NEAR ? SR_F64_NEAR_WAVES2 : SR_F64_WAVES2)
                                      : SR_F64_WAVES)
      : (SWEEP == SWEEP_REC) ? ((MODE == SR_MODE_FAR && M == 2)   ? SR_REC_WAVES_FAR2
                                : (MODE == SR_MODE_FAR && M == 4) ? SR_REC_WAVES_FAR4
                                                                  : SR_REC_WAVES)
                             : (M >= 4 ? 2 : SR_DIRECT_WAVES);
};

// PRE: the step data come from the pre-pass records (batch calls); without
// them (the per-particle and no-allocation calls) the stage forms them from u
template <int MODE, int PREC, int SWEEP, int M, bool PRE>
__global__ void __launch_bounds__(kMaxT)
__attribute__((amdgpu_waves_per_eu(WavesPerEU<MODE, PREC, SWEEP, M>::value)))
k_spectrum(const KParams p) {
  static_assert(!(PREC == SR_PREC_F64 && SWEEP == SWEEP_REC), "recurrence sweep is split-only");
  static_assert(PREC == SR_PREC_SPLIT || M <= 2, "the FP64 sweep holds one or two omegas per thread");
  using Amp = typename std::conditional<PREC == SR_PREC_SPLIT, float, double>::type;
  constexpr bool kSplit = (PREC == SR_PREC_SPLIT);
  constexpr bool kRec = (SWEEP == SWEEP_REC);
  // amplitude words per step.  Far: A is perpendicular to n
  // (n.A = c1 (1 - n.bbar) - c2 n.a = 0 with c1 = (n.a) c2^2), so it is kept
  // as its two components on e_theta, e_phi (the reference sums A_x, A_y, A_z;
  // the two agree to rounding).  Near: v1 = (b - n)/R and v2 = n/R^2.
  constexpr int NA = amp_words<MODE, PREC>();
  constexpr int NC = (MODE == SR_MODE_FAR) ? 2 * NA : 6;   // real accumulators per omega
  constexpr int NP = (MODE == SR_MODE_FAR) ? NC : 12;      // split register partials per omega
  constexpr bool kCompact = kSplit && kRec && (SR_COMPACT != 0);
  // (the direct sweep has no step sort whose barriers it could share: it
  // recomputes tau_{s-1}, which measured 0.5 % faster there)
  constexpr bool kExch = kSplit && kRec && (SR_TAU_EXCHANGE != 0);
  constexpr int kMaxW = kMaxT / kWave;                 // waves per block
  constexpr int kMaxRuns = kMaxW * 4;                  // 64-omega runs per block (M <= 4)

  extern __shared__ __align__(16) char s_dyn_raw[];
  // compaction: smallest |omega|/2pi of each run, its rank, the per-wave
  // bucket histogram, its scan, and the step count of each rank
  __shared__ double s_wmin[kCompact ? kMaxRuns : 1];
  __shared__ int s_rank[kCompact ? kMaxRuns : 1];
  __shared__ double s_wsort[kCompact ? kMaxRuns : 1];   // s_wmin in rank order (ascending)
  __shared__ int s_hist[kCompact ? kMaxRuns * kMaxW : 1];
  __shared__ int s_off[kCompact ? kMaxRuns * kMaxW : 1];
  __shared__ int s_cntb[kCompact ? kMaxRuns + 1 : 1];
  // FP64 sin/cos table (sincos_table) and the largest |omega| of the grid
  constexpr bool kTab = !kSplit && (SR_F64_TABLE != 0);
  __shared__ double2 s_sct[kTab ? kSinCosTab : 1];
  __shared__ double s_wmax[kTab ? kMaxW : 1];
  __shared__ int s_vote[kMaxW];   // per-wave votes of a block-wide any/all

  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int W = T / kWave;
  const int wv = uni(tid / kWave);   // wave-uniform: its derived indices stay scalar
  const int lane = tid % kWave;
  // FP64: a tile whose |tau| * om_abs_max reaches kTabMaxPhase / 2 takes the
  // sin/cos table's wide-phase form
  double om_abs_max = 0.0;
  if (kTab) {
    for (int i = tid; i < kSinCosTab; i += T) {
      double sv, cv;
      sincospi((double)i / (kSinCosTab / 2), &sv, &cv);
      s_sct[i] = make_double2(sv, cv);
    }
    double m = 0.0;
    for (int64_t i = tid; i < (int64_t)p.n_omega; i += T) m = fmax(m, fabs(p.omega[i]));
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (lane == 0) s_wmax[wv] = m;
    __syncthreads();
    for (int w = 0; w < W; w++) om_abs_max = fmax(om_abs_max, s_wmax[w]);
  }
  // dynamic LDS: the staged tile (recurrence: s_rec[T]; else s_tau[T],
  // s_thr[T] (split), s_amp[NA][T]), then the Kahan state, then the particle sums
  static_assert(!kRec || sizeof(RecStep<NA>) == 2 * sizeof(double) + (2 + NA) * sizeof(float),
                "a recurrence record holds exactly the staged step's bytes");
  RecStep<NA> *const s_rec = (RecStep<NA> *)s_dyn_raw;
  double *const s_tau = (double *)s_dyn_raw;
  double *const s_thr = s_tau + T;
  const LdsRows<Amp> s_amp{(Amp *)(s_tau + (kSplit ? 2 : 1) * T), T};
  float *const s_dyn = (float *)(s_dyn_raw + stage_lds_bytes<MODE, PREC, SWEEP>(T));
  float *const s_ks = s_dyn + tid;                       // S[c*M+k] at s_ks[(c*M+k)*T]
  float *const s_kc = s_dyn + (size_t)NC * M * T + tid;  // C likewise
  // particle sums of the single output, k-th omega at s_acc[k*T]
  double *const s_acc = (double *)((char *)s_dyn + (kSplit ? kahan_lds_bytes<MODE, M>(T) : 0)) + tid;
  const double dt = p.dt;
  const double dtInv = 1.0 / dt;
  const bool spheric = (p.comp == SR_COMP_SPHERIC || p.comp == SR_COMP_SPHERIC_COMPLEX);
  const double rot_rev = (double)T * p.omega_step * kInv2Pi;   // revolutions of the k -> k+1 step

  for (int64_t ti = blockIdx.x; ti < p.n_tiles; ti += gridDim.x) {
    const uint32_t tile = p.tile_order ? (uint32_t)p.tile_order[ti] : (uint32_t)ti;
    const uint32_t rest = udiv(tile, p.div_chunks);
    const uint32_t chunk = tile - rest * p.n_chunks;
    const uint32_t qs = udiv(rest, p.div_dir);
    const uint32_t dir = rest - qs * p.n_dir;
    const int q = (int)qs;
    const uint32_t iPhi = udiv(dir, p.div_theta);
    const uint32_t iTh = dir - iPhi * p.n_theta;

    // geometry of the block's point (kernel_farfield.cl:40-42, 384-387;
    // kernel_nearfield.cl:40-42): far n, e_theta, e_phi; near the screen point
    // P.  Formed where it is used (from the block's uniform axis values, scalar
    // loads) rather than held in registers through the sweep.
    // (near: the screen point is held, which measured fewer spilled registers
    // in the FP64 near-field sweeps)
    struct Geom {
      double gx, gy, gz, thx, thy, thz, phx, phy;
    };
    double ngx = 0.0, ngy = 0.0, ngz = 0.0;
    if (MODE == SR_MODE_NEAR) {
      // the block's screen point is wave-uniform: kept in scalar registers
      const double r = p.radius[iTh];
      if (p.screen_y) {
        ngx = unid(r);
        ngy = unid(p.screen_y[iPhi]);
      } else {
        ngx = unid(r * p.cos_p[iPhi]);
        ngy = unid(r * p.sin_p[iPhi]);
      }
      ngz = p.l_screen;
    }
    auto geom = [&]() {
      Geom q;
      if (MODE == SR_MODE_FAR) {
        const double st = p.sin_t[iTh], ct = p.cos_t[iTh];
        const double sp = p.sin_p[iPhi], cp = p.cos_p[iPhi];
        q.gx = st * cp; q.gy = st * sp; q.gz = ct;
        q.thx = ct * cp; q.thy = ct * sp; q.thz = -st;
        q.phx = -sp; q.phy = cp;
      } else {
        q.gx = ngx; q.gy = ngy; q.gz = ngz;
        q.thx = q.thy = q.thz = q.phx = q.phy = 0.0;
      }
      return q;
    };

    // the block's omegas are dealt to its waves in 64-omega runs; the low runs
    // keep the most steps, so the run -> wave assignment is rotated by the tile
    // index to spread the heavy waves of co-resident blocks over the SIMDs
    const int vw = SR_WAVE_ROTATE ? (int)((wv + tile) & (uint32_t)(W - 1)) : wv;   // W is 1, 2 or 4
    // the omega index of the thread's k-th omega: lane of run k*W + vw
    auto omega_slot = [&](int k) { return (int)(chunk * T * M + (k * W + vw) * kWave + lane); };
    int wi[M];
    bool valid[M];
    double om[M];
#pragma unroll
    for (int k = 0; k < M; k++) {
      wi[k] = omega_slot(k);
      valid[k] = wi[k] < (int)p.n_omega;
      om[k] = valid[k] ? p.omega[wi[k]] : 0.0;
    }
    double om_rev[M];
    float omf[M];   // omega of the near field's v1 term (far: unused)
#pragma unroll
    for (int k = 0; k < M; k++) {
      om_rev[k] = om[k] * kInv2Pi;
      omf[k] = (float)om[k];
    }
    // rec sweep: the thread's base frequency in revolutions, the block's omega range
    const double om0_rev = om_rev[0];
    // range of |omega|/2pi over the block (linear grids: the run's end points)
    double w_lo = 0.0, w_hi = 0.0;
    bool rec_sortable = false;
    bool tile_wide = false;   // FP64: the tile's phases may reach kTabMaxPhase / 2
    if (kRec) {
      const int64_t first = (int64_t)chunk * T * M;
      const int64_t last = min((int64_t)(chunk + 1) * T * M, (int64_t)p.n_omega) - 1;
      const double a = p.omega[first] * kInv2Pi, b = p.omega[last] * kInv2Pi;
      w_lo = (a * b <= 0.0) ? 0.0 : fmin(fabs(a), fabs(b));
      w_hi = fmax(fabs(a), fabs(b));
      rec_sortable = p.omega_step > 0.0 && a >= 0.0;
    }
    // compaction: run r = k*W + wave holds the k-th omegas of the wave's lanes;
    // a step keeps some omega of run r iff s_wmin[r] < thr_s.  The recurrence
    // sweep needs the thread's runs in ascending |omega| (a non-negative,
    // increasing grid): then the runs a step keeps are a prefix k < K.
    if (kCompact) {
      __syncthreads();  // an earlier tile of this block is done with s_wmin, s_rank
#pragma unroll
      for (int k = 0; k < M; k++) {
        double v = valid[k] ? fabs(om_rev[k]) : INFINITY;
        for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
        if (lane == 0) s_wmin[k * W + vw] = v;
      }
      __syncthreads();
      if (tid < W * M) {
        const double mine = s_wmin[tid];
        int rk = 0;
        for (int r = 0; r < W * M; r++) {
          const double o = s_wmin[r];
          rk += (o < mine || (o == mine && r < tid)) ? 1 : 0;
        }
        s_rank[tid] = rk;
        s_wsort[rk] = mine;
      }
      // published by the barrier that opens the first step tile
    }
    // output c of this tile (slices: the slice's partial spectra), formed where it is written
    auto dst = [&](int c) -> double * {
      return (p.slices > 1 && p.part_stride == 0)
                 ? p.scratch + ((int64_t)q * p.n_out + c) * p.n_grid * p.n_snaps
                 : p.out[c];
    };

    if (late_s(p.acc_lds)) {
#pragma unroll
      for (int k = 0; k < M; k++) s_acc[k * T] = 0.0;
    }
    const int64_t p_lo = p.slice_lo[q], p_hi = p.slice_lo[q + 1];
    for (int64_t ip = p_lo; ip < p_hi; ip++) {
      const Particle pt = load_particle(p, ip);
      if (pt.n <= 0) continue;
      const int64_t n_act = min(pt.n - 1, (int64_t)pt.ite - 1);
      const double wpdt2 = pt.w * dt * dt;
      const double wpdt = sqrt(pt.w) * dt;

      SnapCursor cur;
      cur.init(pt, p.n_snaps);

      // F64: running sums fs; split: segment partials ps (registers), Kahan
      // sum and compensation in LDS
      double prev[M];
      double fs[M][NC];
      float ps[M][NP];
#pragma unroll
      for (int k = 0; k < M; k++) {
        prev[k] = 0.0;
#pragma unroll
        for (int c = 0; c < NC; c++) fs[k][c] = 0.0;
#pragma unroll
        for (int c = 0; c < NP; c++) ps[k][c] = 0.f;
      }
      if (kSplit) {
#pragma unroll
        for (int i = 0; i < NC * M; i++) {
          s_ks[i * T] = 0.f;
          s_kc[i * T] = 0.f;
        }
      }

      // adds this particle's snapshot value into the outputs the thread owns.
      // Records are rare; the split sweeps walk their omegas one at a time
      // (not unrolled, everything re-derived from k) so that the record code
      // does not set the kernel's register budget.
      // one omega's record: F are its sums (FP64), omk its 2 pi omega
      auto emit_k = [&](uint32_t slot, int k, int wk, const double *F, double omk) {
        double re[3], im[3];
        if constexpr (MODE == SR_MODE_FAR) {
          // F = F_theta e_theta + F_phi e_phi (spheric, kernel_farfield.cl:442-445:
          // the radial part n.F is 0, as n.A is)
          if (spheric) {
            re[0] = 0.0; re[1] = F[0]; re[2] = F[1];
            im[0] = 0.0; im[1] = F[2]; im[2] = F[3];
          } else {
            const Geom q = geom();
            re[0] = F[0] * q.thx + F[1] * q.phx; re[1] = F[0] * q.thy + F[1] * q.phy; re[2] = F[0] * q.thz;
            im[0] = F[2] * q.thx + F[3] * q.phx; im[1] = F[2] * q.thy + F[3] * q.phy; im[2] = F[2] * q.thz;
          }
        } else {
          // near: the sums are the field, [Re F (3) | Im F (3)]
          (void)omk;
#pragma unroll
          for (int c = 0; c < 3; c++) {
            re[c] = F[c];
            im[c] = F[3 + c];
          }
        }
        double v[kMaxOut];
        switch (p.comp) {
          case kCompField:
#pragma unroll
            for (int c = 0; c < 3; c++) {
              v[2 * c] = re[c];
              v[2 * c + 1] = im[c];
            }
            break;
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
            const double ff = (MODE == SR_MODE_FAR && p.comp == SR_COMP_CARTESIAN_COMPLEX)
                                  ? p.form_factor[wk] : 1.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
              v[2 * c] = wpdt * (ff * re[c]);
              v[2 * c + 1] = wpdt * (ff * im[c]);
            }
          }
        }
        if (late_s(p.acc_lds)) {
          s_acc[k * T] += v[0];
        } else {
          const int64_t g = (int64_t)slot * p.n_grid + (int64_t)dir * p.n_omega + late(wk) +
                            (ip - p.part_base) * p.part_stride;
#pragma unroll
          for (int c = 0; c < kMaxOut; c++)
            if (c < p.n_out) dst(c)[g] += v[c];
        }
      };
      auto emit = [&](uint32_t slot) {
        if constexpr (kSplit) {
#pragma unroll 1
          for (int k = 0; k < M; k++) {
            const int wk = omega_slot(k);
            if (wk >= (int)p.n_omega) continue;
            double F[NC];
#pragma unroll
            for (int c = 0; c < NC; c++)
              F[c] = (double)s_ks[(c * M + k) * T] - (double)s_kc[(c * M + k) * T];
            emit_k(slot, k, wk, F, p.omega[wk]);
          }
        } else {
          // FP64: the sums live in registers, so the loop is unrolled
#pragma unroll
          for (int k = 0; k < M; k++)
            if (valid[k]) emit_k(slot, k, wi[k], fs[k], om[k]);
        }
      };

      // one step of the recurrence sweep for the thread's first K omegas: the
      // base phase in FP64, then each omega by rotation.  MASK 0: no masking
      // (tiles where no step cuts the block's omega range); 1: the step is
      // skipped when the whole block is above its threshold and each omega is
      // masked otherwise; 2: only the K-th omega is masked (sorted tiles, where
      // the runs below the step's last kept run keep all their omegas).
      // rec_acc: the accumulations of one step from its base sin/cos
      // The second omega comes from the first by one complex multiplication;
      // the third and fourth by the three-term recurrence z_{k+1} = 2 cos(d)
      // z_k - z_{k-1} of equally spaced angles (two FMAs instead of two
      // multiplications and two FMAs; over two terms its error growth is
      // that of the multiplication's within a factor 3)
      auto rec_acc = [&](auto Kc, auto Mc, float sn, float cn, double thr, float er, float ei,
                         const float *a) {
        constexpr int K = decltype(Kc)::value;
        constexpr int MASK = decltype(Mc)::value;
        float cp = 0.f, sp = 0.f;   // z_{k-1}
        const float c2 = (SR_REC_CHEB && K >= 3) ? er + er : 0.f;
#pragma unroll
        for (int k = 0; k < K; k++) {
          float cm = cn, sm = sn;
          if (MASK == 1 || (MASK == 2 && k == K - 1)) {
            const bool act = fabs(om_rev[k]) < thr;
            cm = act ? cn : 0.f;
            sm = act ? sn : 0.f;
          }
          split_acc<MODE, NA>(ps[k], a, cm, sm);
          if (k + 1 < K) {
            float cn2, sn2;
            if (SR_REC_CHEB && k >= 1) {
              cn2 = __builtin_fmaf(c2, cn, -cp);
              sn2 = __builtin_fmaf(c2, sn, -sp);
            } else {
              cn2 = __builtin_fmaf(cn, er, -sn * ei);
              sn2 = __builtin_fmaf(cn, ei, sn * er);
            }
            cp = cn;
            sp = sn;
            cn = cn2;
            sn = sn2;
          }
        }
      };
      auto rec_step = [&](auto Kc, auto Mc, double tau, double thr, float er, float ei, const float *a) {
        if (decltype(Mc)::value == 1 && thr <= w_lo) return;   // every omega of the block filtered out
        float sn, cn;
        sincos_rev(om0_rev * tau, &sn, &cn);
        rec_acc(Kc, Mc, sn, cn, thr, er, ei, a);
      };
      // staged step jj of the recurrence sweep: two (far) or three (near)
      // 16-byte reads of its record
      auto rec_load = [&](int jj, double &tau, double &thr, float &er, float &ei, float *a) {
        const RecStep<NA> &r = s_rec[jj];
        const double2 tt = *(const double2 *)&r.tau;
        const float4 u = *(const float4 *)&r.er;
        tau = tt.x;
        thr = tt.y;
        er = u.x;
        ei = u.y;
        a[0] = u.z;
        a[1] = u.w;
        if constexpr (NA > 2) {
          const float4 w = *(const float4 *)&r.amp[2];
          a[2] = w.x;
          a[3] = w.y;
          a[4] = w.z;
          a[5] = w.w;
        }
      };
      // steps [j, jend) of the staged tile: groups of 4, then the tail
      auto rec_sweep = [&](auto Kc, auto Mc, int &j, int jend) {
        for (; j + 4 <= jend; j += 4) {
          double tq[4], hq[4];
          float erq[4], eiq[4], aq[4][NA];
#pragma unroll
          for (int q = 0; q < 4; q++) rec_load(j + q, tq[q], hq[q], erq[q], eiq[q], aq[q]);
          if (SR_REC_PIPE && decltype(Mc)::value != 1) {
            // the four base phases and sin/cos first (independent chains), then
            // the accumulations
            float sn[4], cn[4];
#pragma unroll
            for (int q = 0; q < 4; q++) sincos_rev(om0_rev * tq[q], &sn[q], &cn[q]);
#pragma unroll
            for (int q = 0; q < 4; q++) rec_acc(Kc, Mc, sn[q], cn[q], hq[q], erq[q], eiq[q], aq[q]);
          } else {
#pragma unroll
            for (int q = 0; q < 4; q++) rec_step(Kc, Mc, tq[q], hq[q], erq[q], eiq[q], aq[q]);
          }
        }
        for (; j < jend; j++) {
          double tau, thr;
          float er, ei, a[NA];
          rec_load(j, tau, thr, er, ei, a);
          rec_step(Kc, Mc, tau, thr, er, ei, a);
        }
      };

      // one staged step of the direct sweeps; WIDE (FP64): the tile's phases
      // may reach the sin/cos table's limit
      auto direct_step = [&](int j, auto Wc) {
        const double tau = s_tau[j];
        const double thr = kSplit ? s_thr[j] : 0.0;
        Amp a[NA];
#pragma unroll
        for (int c = 0; c < NA; c++) a[c] = s_amp[c][j];
#pragma unroll
        for (int k = 0; k < M; k++) {
          // FP64: phase in radians and the reference's exact test
          // |phase_s - phase_{s-1}| < pi.  Split: phase in revolutions
          // (omega/2pi precomputed) and the same inequality in its
          // threshold form |omega|/2pi < thr_s, staged per step; the two
          // can disagree only within rounding of the boundary.
          const double ph = (kSplit ? om_rev[k] : om[k]) * tau;
          bool act;
          if (kSplit) {
            act = valid[k] && (fabs(om_rev[k]) < thr);
          } else {
            const double d = ph - prev[k];
            prev[k] = ph;
            act = valid[k] && (fabs(d) < kPi);
          }
          if (act) {
            if (kSplit) {
              float sn, cn;
              sincos_rev(ph, &sn, &cn);
              float af[NA];
#pragma unroll
              for (int c = 0; c < NA; c++) af[c] = (float)a[c];
              split_acc<MODE, NA>(ps[k], af, cn, sn);
            } else {
              double sn, cn;
              if (kTab)
                sincos_table(ph, s_sct, decltype(Wc)::value != 0, &sn, &cn);
              else
                sincos_reduced(ph, &sn, &cn);
              double ad[NA];
#pragma unroll
              for (int c = 0; c < NA; c++) ad[c] = (double)a[c];
              accumulate<MODE, NA>(fs[k], ad, cn, sn, om[k]);
            }
          }
        }
      };
      // steps [j, jend) of the direct sweep.  Split: single steps up to a
      // multiple of 4, then groups of 4 read with 16-byte LDS reads (each
      // frequency still skipped where the filter rejects it), then the tail
      auto direct_sweep = [&](int &j, int jend) {
        if constexpr (kSplit && SR_DIRECT_GROUP) {
          const int jg = min(jend, (j + 3) & ~3);
          for (; j < jg; j++) direct_step(j, IC<0>());
          for (; j + 4 <= jend; j += 4) {
            const double2 t01 = *(const double2 *)&s_tau[j];
            const double2 t23 = *(const double2 *)&s_tau[j + 2];
            const double2 h01 = *(const double2 *)&s_thr[j];
            const double2 h23 = *(const double2 *)&s_thr[j + 2];
            float4 av[NA];
#pragma unroll
            for (int c = 0; c < NA; c++) av[c] = *(const float4 *)&s_amp[c][j];
            const double tq[4] = {t01.x, t01.y, t23.x, t23.y};
            const double hq[4] = {h01.x, h01.y, h23.x, h23.y};
#pragma unroll
            for (int q = 0; q < 4; q++) {
              float a[NA];
#pragma unroll
              for (int c = 0; c < NA; c++) a[c] = lane4(av[c], q);
#pragma unroll
              for (int k = 0; k < M; k++) {
                if (valid[k] && fabs(om_rev[k]) < hq[q]) {
                  float sn, cn;
                  sincos_rev(om_rev[k] * tq[q], &sn, &cn);
                  split_acc<MODE, NA>(ps[k], a, cn, sn);
                }
              }
            }
          }
        }
        if (kTab && tile_wide) {
          for (; j < jend; j++) direct_step(j, IC<1>());
          return;
        }
#pragma unroll SR_F64_UNROLL
        for (; j < jend; j++) direct_step(j, IC<0>());
      };
      // split: Kahan fold of the segment's partial sums into the LDS S, C
      auto kahan_fold = [&]() {
        if (!kSplit) return;
#pragma unroll
        for (int k = 0; k < M; k++) {
          float part[NC];
#pragma unroll
          for (int c = 0; c < NC; c++) part[c] = ps[k][c];
          if constexpr (MODE != SR_MODE_FAR) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
              part[c] = fmx(-omf[k], ps[k][6 + c], part[c]);
              part[3 + c] = fmx(omf[k], ps[k][9 + c], part[3 + c]);
            }
          }
#pragma unroll
          for (int c = 0; c < NC; c++) {
            float &S = s_ks[(c * M + k) * T];
            float &C = s_kc[(c * M + k) * T];
            const float yv = part[c] - C;
            const float tv = S + yv;
            C = (tv - S) - yv;
            S = tv;
          }
#pragma unroll
          for (int c = 0; c < NP; c++) ps[k][c] = 0.f;
        }
      };

      for (int64_t t0 = 0; t0 < n_act; t0 += T) {
        const int cnt = uni((int)min((int64_t)T, n_act - t0));
        __syncthreads();  // previous tile fully consumed
        int step_on = 1;
        int st_wide = 0;
        Amp st_amp[NA];
        double st_tau = 0.0, st_thr = 0.0, st_prev = 0.0;
        float st_rot[2] = {1.f, 0.f};
        if (tid < cnt) {
          // ---- stage: omega-independent part of step s = t0 + tid, in FP64 ----
          const int64_t s = t0 + tid;
          const int64_t gs = pt.off + s;
          const double time = (double)(uint32_t)(pt.its + (uint32_t)s) * dt;
          const double xs = p.x[gs], ys = p.y[gs], zs = p.z[gs];
          const Geom q = geom();
          const double gx = q.gx, gy = q.gy, gz = q.gz;
          const double thx = q.thx, thy = q.thy, thz = q.thz, phx = q.phx, phy = q.phy;
          double tau;
          // tau_{s-1} (phase_{-1} = 0 at the first step): recomputed here by the
          // tile's first thread, or by every thread without the exchange
          const bool own_prev = kSplit && s > 0 && (!kExch || tid == 0);
          if (MODE == SR_MODE_FAR) {
            // kernel_farfield.cl:65-94
            tau = time - (xs * gx + ys * gy + zs * gz);
            if (own_prev) {
              const double tp = (double)(uint32_t)(pt.its + (uint32_t)(s - 1)) * dt;
              st_prev = tp - (p.x[gs - 1] * gx + p.y[gs - 1] * gy + p.z[gs - 1] * gz);
            }
            double ax, ay, az, mx, my, mz;
            if constexpr (PRE) {
              const double2 *r = (const double2 *)(p.pre + 6 * gs);
              const double2 r0 = r[0], r1 = r[1], r2 = r[2];
              ax = r0.x; ay = r0.y; az = r1.x;
              mx = r1.y; my = r2.x; mz = r2.y;
            } else {
              step_am(p.ux, p.uy, p.uz, gs, dtInv, &ax, &ay, &az, &mx, &my, &mz);
            }
            double c1 = ax * gx + ay * gy + az * gz;
            double c2 = 1.0 - (mx * gx + my * gy + mz * gz);
            // split: the amplitude ends in FP32, so the hardware reciprocal
            // (relative error <= 4.6e-8 measured on gfx950) with one Newton
            // step (~1e-15) replaces the IEEE division
            if (kSplit && SR_SPLIT_RCP) {
              const double r = __builtin_amdgcn_rcp(c2);
              c2 = fma(r, fma(-c2, r, 1.0), r);
            } else {
              c2 = 1.0 / c2;
            }
            c1 = c1 * c2 * c2;
            const double Ax = c1 * (gx - mx) - c2 * ax;
            const double Ay = c1 * (gy - my) - c2 * ay;
            const double Az = c1 * (gz - mz) - c2 * az;
            // the two transverse components; the record rebuilds x, y, z
            st_amp[0] = (Amp)(thx * Ax + thy * Ay + thz * Az);
            st_amp[1] = (Amp)(phx * Ax + phy * Ay);
          } else {
            // kernel_nearfield.cl:67-88; omega of c1 = omega v1 is applied in the sums
            double b0x, b0y, b0z;
            if constexpr (PRE) {
              const double *r = p.pre + 3 * gs;
              b0x = r[0]; b0y = r[1]; b0z = r[2];
            } else {
              const double g0 = gamma_inv(p.ux[gs], p.uy[gs], p.uz[gs]);
              b0x = p.ux[gs] * g0; b0y = p.uy[gs] * g0; b0z = p.uz[gs] * g0;
            }
            const double rx = gx - xs, ry = gy - ys, rz = gz - zs;
            const double R = sqrt(rx * rx + ry * ry + rz * rz);
            tau = time + R;
            if (own_prev) {
              const double tp = (double)(uint32_t)(pt.its + (uint32_t)(s - 1)) * dt;
              const double qx = gx - p.x[gs - 1], qy = gy - p.y[gs - 1], qz = gz - p.z[gs - 1];
              st_prev = tp + sqrt(qx * qx + qy * qy + qz * qz);
            }
            const double rInv = 1.0 / R;
            const double nx = rInv * rx, ny = rInv * ry, nz = rInv * rz;
            const double r2 = rInv * rInv;
            st_amp[0] = (Amp)(rInv * (b0x - nx));
            st_amp[1] = (Amp)(rInv * (b0y - ny));
            st_amp[2] = (Amp)(rInv * (b0z - nz));
            st_amp[3] = (Amp)(r2 * nx);
            st_amp[4] = (Amp)(r2 * ny);
            st_amp[5] = (Amp)(r2 * nz);
          }
          st_tau = tau;
          if (kTab) st_wide = fabs(tau) * om_abs_max >= 0.5 * kTabMaxPhase;
          if (kRec) sincos_rev(rot_rev * tau, &st_rot[1], &st_rot[0]);
          if constexpr (kExch) s_rec[tid].tau = tau;   // in step order, for the next thread
        }
        if constexpr (kSplit) {
          if (kExch) {
            // every tau of the tile is in LDS; the staged copies are written
            // after the step sort's barriers (or, unsorted, over the same value)
            __syncthreads();
            if (tid > 0 && tid < cnt) st_prev = s_rec[tid - 1].tau;
          }
          if (tid < cnt) {
            // |omega tau_s - omega tau_{s-1}| < pi  <=>  |omega|/2pi < (1/2)/|tau_s - tau_{s-1}|
            // (phase_{-1} = 0 at the first step); 1/0 gives +inf: always active
            // (split: the hardware reciprocal with one Newton step, as for c2;
            // the raw one, 4.6e-8, moved enough boundary decisions to raise the
            // largest element error of the betatron spectrum from 1e-7 to 6e-7)
            // (d = 0: +inf, always active, as the division gives; the Newton
            // step alone would make it fma(-0, inf, 1) = NaN)
            double thr;
            if (SR_SPLIT_RCP) {
              const double d = fabs(st_tau - st_prev);
              const double r = __builtin_amdgcn_rcp(d);
              thr = (d == 0.0) ? INFINITY : 0.5 * fma(r, fma(-d, r, 1.0), r);
            } else {
              thr = 0.5 / fabs(st_tau - st_prev);
            }
            st_thr = thr;
            step_on = thr > w_hi;
          }
        }
        // compaction (tiles without a snapshot record): bucket_s = number of runs
        // that keep some omega at step s; a stable counting sort puts the steps
        // in descending bucket order, so run r of rank rho sweeps exactly the
        // first s_cntb[rho] steps.  Steps of bucket 0 are needed by no run.
        const bool sort_tile = kCompact && rec_sortable && uni((cur.next >= 0 && cur.next < t0 + cnt) ? 0 : 1);
        int pos = tid;
        if (sort_tile) {
          const int R = W * M;
          int bucket = 0;
          if (tid < cnt) {
            if (SR_BUCKET_SEARCH) {
              // R is a power of two; the count of sorted entries below thr
              for (int h = R >> 1; h > 0; h >>= 1)
                if (s_wsort[bucket + h - 1] < st_thr) bucket += h;
              if (s_wsort[bucket] < st_thr) bucket++;
            } else {
              for (int r = 0; r < R; r++) bucket += (s_wmin[r] < st_thr) ? 1 : 0;
            }
          }
          // per-wave counts in the order (bucket R..1) x (wave 0..W-1)
          uint64_t mine = 0;
          for (int b = 1; b <= R; b++) {
            const uint64_t m = __ballot(bucket == b);
            if (bucket == b) mine = m;
            if (lane == 0) s_hist[(R - b) * W + wv] = __popcll(m);
          }
          __syncthreads();
          if (wv == 0) {
            const int n_ent = R * W;   // <= 64: one entry per lane
            const int v = lane < n_ent ? s_hist[lane] : 0;
            int inc = v;
            for (int o = 1; o < kWave; o <<= 1) {
              const int y = __shfl_up(inc, o);
              if (lane >= o) inc += y;
            }
            if (lane < n_ent) s_off[lane] = inc - v;
            // steps with bucket > b = everything before entry (b, wave 0)
            if (lane < n_ent && lane % W == 0) s_cntb[R - lane / W] = inc - v;
            if (lane == n_ent - 1) s_cntb[0] = inc;
          }
          __syncthreads();
          const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
          pos = (bucket > 0) ? s_off[(R - bucket) * W + wv] + __popcll(mine & below) : -1;
        }
        if (tid < cnt && pos >= 0) {
          pos = late(pos);   // the staged slots' LDS addresses formed here
          if constexpr (kRec) {
            RecStep<NA> &r = s_rec[pos];
            *(double2 *)&r.tau = make_double2(st_tau, st_thr);
            *(float4 *)&r.er = make_float4(st_rot[0], st_rot[1], (float)st_amp[0], (float)st_amp[1]);
            if constexpr (NA > 2)
              *(float4 *)&r.amp[2] = make_float4((float)st_amp[2], (float)st_amp[3], (float)st_amp[4], (float)st_amp[5]);
          } else {
            s_tau[pos] = st_tau;
            if (kSplit) s_thr[pos] = st_thr;
#pragma unroll
            for (int c = 0; c < NA; c++) s_amp[c][pos] = st_amp[c];
          }
        }
        // rec: the tile needs no masking when every step keeps the whole block.
        // The block votes through one ballot per wave and an LDS word per wave
        // (the block is one-dimensional: unlike the library's work-group
        // reductions this keeps no y/z work-item ids alive, which cost the FP64
        // near-field variant two spilled registers)
        bool tile_on = false;
        if constexpr (kRec || kTab) {
          const bool pred = kRec ? (step_on == 0) : (st_wide != 0);   // rec: any step not keeping all
          const uint64_t b = __ballot(pred);
          if (lane == 0) s_vote[wv] = (b != 0) ? 1 : 0;
          __syncthreads();
          int any = 0;
          for (int w = 0; w < W; w++) any |= s_vote[w];
          if constexpr (kRec)
            tile_on = uni(any) == 0;
          else
            tile_wide = uni(any) != 0;
        } else {
          __syncthreads();
        }

        // ---- sweep, split at snapshot records falling inside the tile ----
        if (sort_tile) {
          // run k of this wave (rank s_rank[k*W + wave]) needs the first ck[k]
          // steps; the runs ascend, so ck[0] >= ck[1] >= ... and the steps
          // [ck[K], ck[K-1]) keep exactly the runs k < K (ck[M] = 0)
          int ck[M];
#pragma unroll
          for (int k = 0; k < M; k++) ck[k] = uni(s_cntb[s_rank[k * W + vw]]);
          int j = 0;
          if (tile_on) {
            // every valid omega of the block kept at every step: the wave's
            // valid runs all have ck = cnt, runs past n_omega have ck = 0
            rec_sweep(IC<M>(), IC<0>(), j, ck[0]);
          } else {
            rec_sweep(IC<M>(), IC<2>(), j, ck[M - 1]);
            if (M >= 2) rec_sweep(IC<(M >= 2 ? M - 1 : 1)>(), IC<2>(), j, ck[M >= 2 ? M - 2 : 0]);
            if (M >= 3) rec_sweep(IC<(M >= 3 ? M - 2 : 1)>(), IC<2>(), j, ck[M >= 3 ? M - 3 : 0]);
            if (M >= 4) rec_sweep(IC<(M >= 4 ? M - 3 : 1)>(), IC<2>(), j, ck[0]);
          }
          kahan_fold();
        } else {
          int j = 0;
          while (j < cnt) {
            const bool rec_here = (cur.next >= 0) && (cur.next < t0 + cnt);
            // records are strictly increasing, so cur.next >= t0 + j; the max()
            // only guarantees forward progress
            const int jend = uni(rec_here ? max(j + 1, (int)(cur.next - t0) + 1) : cnt);
            if (kRec) {
              if (tile_on)
                rec_sweep(IC<M>(), IC<0>(), j, jend);
              else
                rec_sweep(IC<M>(), IC<1>(), j, jend);
            } else {
              direct_sweep(j, jend);
            }
            kahan_fold();
            if (rec_here) {
              emit(cur.slot);
              const int64_t last = cur.next;
              cur.slot++;
              cur.advance(pt, p.n_snaps, last);
            }
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
    if (late_s(p.acc_lds)) {
#pragma unroll 1
      for (int k = 0; k < M; k++) {
        const int wk = omega_slot(k);
        if (wk < (int)p.n_omega) dst(0)[(int64_t)dir * p.n_omega + wk] += s_acc[k * T];
      }
    }
  }
}


// ---------------------------------------------------------------------------
// heaviest tiles first
// ---------------------------------------------------------------------------
// A far-field tile's time grows with the share of its (omega, step) pairs that
// the filter keeps, which varies several-fold over the directions (near the
// beam axis most pairs are kept).  Launched in index order, the last blocks
// keep a few CUs busy while the rest of the GPU idles; so each tile's time is
// estimated from a sample of steps and the tiles are launched heaviest first.
// Every tile writes only its own outputs, so the order changes no result.
#ifndef SR_ORDER_SAMPLES
#define SR_ORDER_SAMPLES 256
#endif
#ifndef SR_ORDER_STEP_COST
#define SR_ORDER_STEP_COST 0.1f
#endif
constexpr int kOrderSamples = SR_ORDER_SAMPLES;
constexpr int64_t kOrderMinTiles = 2048;
constexpr int64_t kOrderMaxTiles = 65536;
constexpr double kOrderMinEvals = 1e10;
// the ranking compares every pair of tiles (2.3 ms for 32768 tiles); it is
// used when the launch has at least kOrderEvalsPerTile2 * tiles^2 evaluations
// (about 100 times the ranking's time at 5e12 evaluations/s)
constexpr double kOrderEvalsPerTile2 = 1.2e3;
// launches from this size on are planned with more particle slices (finer tiles)
constexpr double kOrderSliceEvals = 1e12;
constexpr float kOrderStepCost = SR_ORDER_STEP_COST;   // a step's fixed work, in units of a fully kept step's sweep

// cost[tile] for every (direction, omega chunk), the same for each particle
// slice: sum over sampled intervals of kOrderStepCost + the share of the
// chunk's |omega|/2pi range below the interval's threshold 1/(2|dtau|)
__global__ void k_tile_cost(const KParams p, int64_t run_len, float *cost) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)p.n_dir * p.n_chunks) return;
  const uint32_t chunk = (uint32_t)(i % p.n_chunks);
  const uint32_t dir = (uint32_t)(i / p.n_chunks);
  const uint32_t iPhi = dir / p.n_theta;
  const uint32_t iTh = dir - iPhi * p.n_theta;
  const double st = p.sin_t[iTh], ct = p.cos_t[iTh];
  const double nx = st * p.cos_p[iPhi], ny = st * p.sin_p[iPhi], nz = ct;
  const int64_t first = (int64_t)chunk * run_len;
  const int64_t last = min(first + run_len, (int64_t)p.n_omega);
  double wlo = 1e300, whi = 0.0;
  for (int64_t w = first; w < last; w++) {
    const double a = fabs(p.omega[w]) * kInv2Pi;
    wlo = fmin(wlo, a);
    whi = fmax(whi, a);
  }
  float c = 0.f;
  const int64_t n = p.n_points;
  for (int k = 0; k < kOrderSamples; k++) {
    const int64_t g = 1 + (int64_t)k * (n - 1) / kOrderSamples;   // intervals (g-1, g); rarely across particles
    const double dtau = p.dt - (nx * (p.x[g] - p.x[g - 1]) + ny * (p.y[g] - p.y[g - 1]) + nz * (p.z[g] - p.z[g - 1]));
    const double thr = 0.5 / fabs(dtau);
    double f = (whi > wlo) ? (thr - wlo) / (whi - wlo) : (thr > wlo ? 1.0 : 0.0);
    f = fmin(fmax(f, 0.0), 1.0);
    c += kOrderStepCost + (float)f;
  }
  if (!(c >= 0.f && c <= 1e30f)) c = 0.f;   // a strict order below needs no NaN
  for (int q = 0; q < p.slices; q++) cost[chunk + (int64_t)p.n_chunks * (dir + (int64_t)p.n_dir * q)] = c;
}

// order[rank] = tile, rank = tiles that cost more, or as much with a smaller
// index: a permutation (the key (-cost, index) is a strict total order).
// (Keeping the particle slices in launch order and sorting only within each,
// for cache locality, measured 10 % slower on the 20000-particle batch.)
__global__ void __launch_bounds__(256) k_tile_rank(const float *cost, int64_t n, int32_t *order) {
  __shared__ float s_c[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const float ci = i < n ? cost[i] : 0.f;
  int64_t rank = 0;
  for (int64_t j0 = 0; j0 < n; j0 += 256) {
    __syncthreads();
    // past n: -1, below every cost, so never counted
    s_c[threadIdx.x] = (j0 + threadIdx.x < n) ? cost[j0 + threadIdx.x] : -1.f;
    __syncthreads();
    int cnt = 0;
#pragma unroll 16
    for (int jj = 0; jj < 256; jj++) {
      const float cj = s_c[jj];
      cnt += (cj > ci || (cj == ci && j0 + jj < i)) ? 1 : 0;
    }
    rank += cnt;
  }
  if (i < n) order[rank] = (int32_t)i;
}

// Direction-independent data of every sample (the stage's inputs that do not
// depend on the spectral point), computed once per launch: far a, bbar of the
// interval (i, i+1) (kernel_farfield.cl:74-83); near b_i (kernel_nearfield.cl:
// 79-82).  The expressions are the stage's own, so the results are identical.
template <int MODE>
__global__ void k_step_prep(const double *ux, const double *uy, const double *uz, int64_t n, double dt,
                            double *pre) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (MODE == SR_MODE_FAR) {
    double ax = 0.0, ay = 0.0, az = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    if (i + 1 < n) step_am(ux, uy, uz, i, 1.0 / dt, &ax, &ay, &az, &mx, &my, &mz);
    double2 *r = (double2 *)(pre + 6 * i);
    r[0] = make_double2(ax, ay);
    r[1] = make_double2(az, mx);
    r[2] = make_double2(my, mz);
  } else {
    const double g0 = gamma_inv(ux[i], uy[i], uz[i]);
    double *r = pre + 3 * i;
    r[0] = ux[i] * g0; r[1] = uy[i] * g0; r[2] = uz[i] * g0;
  }
}

// out[c][i] += sum_q scratch[q][c][i], q in order
__global__ void k_reduce_slices(const KParams p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)p.n_out * p.n_grid * p.n_snaps;
  if (i >= total) return;
  const int64_t per = (int64_t)p.n_grid * p.n_snaps;
  const int c = (int)(i / per);
  const int64_t g = i - (int64_t)c * per;
  double s = 0.0;
  for (int q = 0; q < p.slices; q++) s += p.scratch[((int64_t)q * p.n_out + c) * per + g];
  p.out[c][g] += s;
}

__global__ void k_fill_zero(double *buf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// scratch of the no-allocation entry point
// ---------------------------------------------------------------------------
// sr_spectrum_batch_noalloc keeps its workspace (the particle slices' partial
// spectra and the tile launch order) in this array instead of a caller's
// buffer.  Each GPU holds its own copy from the library's load to the end of
// the process.  On the host the name is only a stand-in: its device address
// on the current GPU comes from hipGetSymbolAddress, once per GPU.
constexpr size_t kNoAllocScratchBytes = SR_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current GPU, looked up on the first call
// made with that GPU current
int noalloc_scratch(void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(SR_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SR_OK;
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
  int sweep;     // SWEEP_*
  int waves;     // waves per block
  int M;         // omegas per thread
  int slices;
  int64_t blocks;
};

// omegas per thread each variant is built for (the near-field recurrence
// keeps twice the far field's accumulators, so it stops at 2)
int allowed_m(int mode, int prec, int sweep, int m) {
  if (prec == SR_PREC_F64) return m == 1 || m == 2;
  if (sweep == SWEEP_REC) return mode == SR_MODE_FAR ? (m == 2 || m == 4) : m == 2;
  return m == 1 || m == 2 || (mode == SR_MODE_FAR && m == 4);
}

// Tile shape: prefer the largest omega run per block (T*M) that is mostly
// used, then enough particle slices to put ~8 waves on each of the 1024 SIMDs.
// n_points: samples of the batch (0: unknown), for the size of the launch
Plan make_plan(const sr_grid &g, const sr_opts &o, int64_t n_particles, uint32_t n_snaps, int64_t n_points = 0) {
  Plan pl;
  const int64_t n_om = g.n_omega;
  const int64_t n_dir = (int64_t)g.n_theta * g.n_phi;
  // AUTO: the recurrence sweep for split spectra on linear grids (measured
  // fastest there, far and near), the direct sweep otherwise
  pl.sweep = SWEEP_DIRECT;
  if (o.precision == SR_PREC_SPLIT && g.omega_step != 0.0 &&
      (o.sweep == SR_SWEEP_RECURRENCE || o.sweep == SR_SWEEP_AUTO))
    pl.sweep = SWEEP_REC;
  int best_w = 1, best_m = 1;
  double best_score = -1;
  const int m_opts[3] = {1, 2, 4};
  for (int w = 1; w <= 4; w *= 2) {
    for (int mi = 0; mi < 3; mi++) {
      const int m = m_opts[mi];
      if (!allowed_m(g.mode, o.precision, pl.sweep, m)) continue;
      const int64_t run = (int64_t)kWave * w * m;
      const int64_t chunks = (n_om + run - 1) / run;
      const double eff = (double)n_om / (double)(chunks * run);
      if (eff < 0.75) continue;
      // far-field recurrence: 4 omegas per thread in 2-wave blocks (one
      // 512-omega chunk per direction, 3 waves per SIMD; when chosen: 5934 G
      // evals/s on the betatron beam against 5267 for 2 omegas x 2 waves and
      // 4871 for 2 x 4); elsewhere M = 2 measured fastest.  Blocks
      // of 2 waves: with the tiles launched heaviest first, the finer tiles
      // balance better (+8 % over 4 waves)
      const bool rec_far = pl.sweep == SWEEP_REC && g.mode == SR_MODE_FAR;
      const double mpref = rec_far ? (m == 4 ? 0.04 : (m == 2 ? 0.02 : 0.0)) : ((m == 2) ? 0.02 : 0.0);
      const double wpref = (pl.sweep == SWEEP_REC && g.mode == SR_MODE_FAR && w == 2) ? 0.03 : 0.0;
      // launches too small to fill the GPU (one particle: no slices to add
      // waves) prefer the plan with the most waves: 256 CUs hold ~8K waves
      const double fill =
          std::min(1.0, (double)(chunks * n_dir * w * std::min<int64_t>(std::max<int64_t>(n_particles, 1), kMaxSlices)) /
                            8192.0);
      const double score = eff * fill + 0.01 * w + mpref + wpref;
      if (score > best_score) { best_score = score; best_w = w; best_m = m; }
    }
  }
  if (best_score < 0) {
    best_w = 1;
    best_m = (pl.sweep == SWEEP_REC) ? 2 : 1;
  }
  pl.waves = o.waves_per_block > 0 ? o.waves_per_block : best_w;
  pl.M = o.omega_per_lane > 0 ? o.omega_per_lane : best_m;
  if (!allowed_m(g.mode, o.precision, pl.sweep, pl.M)) pl.M = (pl.sweep == SWEEP_REC) ? 2 : 1;
  const int64_t run = (int64_t)kWave * pl.waves * pl.M;
  const int64_t chunks = (n_om + run - 1) / run;
  const int64_t base_waves = chunks * n_dir * pl.waves;
  int s = 1;
  if (o.slices > 0) {
    s = o.slices;
  } else {
    // particle slices until the launch has ~8K waves; large launches until
    // ~64K: smaller tiles leave a shorter tail (betatron beam +3 %, the uneven
    // 20000-particle batch +20 % at 16 slices, the near-field screen +6 % at
    // 12 of equal-cost tiles).  Far-field tiles run heaviest first, and only
    // as far as they stay orderable: more slices without the order would
    // lengthen the tail instead.
    const double evals = (double)n_points * (double)(n_om * n_dir);
    const bool far_ordered = g.mode == SR_MODE_FAR && o.tile_order == 0;
    const bool big = evals >= kOrderSliceEvals && (g.mode != SR_MODE_FAR || far_ordered);
    const int64_t target = big ? 64 * 1024 : 8 * 1024;
    auto orderable = [&](int64_t tiles) {
      return tiles <= kOrderMaxTiles && evals >= kOrderEvalsPerTile2 * (double)tiles * (double)tiles;
    };
    while (s < kMaxSlices && base_waves * s < target && (int64_t)(s * 2) <= n_particles &&
           (!(big && far_ordered) || orderable(chunks * n_dir * s * 2)))
      s *= 2;
  }
  if (s > n_particles) s = (int)std::max<int64_t>(1, n_particles);
  pl.slices = s;
  pl.blocks = chunks * n_dir * s;
  (void)n_snaps;
  return pl;
}

template <int MODE, int PREC, int SWEEP, int M>
hipError_t launch_t(const KParams &kp, int64_t grid, int threads, hipStream_t st) {
  const size_t lds = stage_lds_bytes<MODE, PREC, SWEEP>(threads) +
                     ((PREC == SR_PREC_SPLIT) ? kahan_lds_bytes<MODE, M>(threads) : 0) +
                     (kp.acc_lds ? acc_lds_bytes<M>(threads) : 0);
  if (kp.pre)
    hipLaunchKernelGGL((k_spectrum<MODE, PREC, SWEEP, M, true>), dim3((unsigned)grid), dim3(threads), lds, st,
                       kp);
  else
    hipLaunchKernelGGL((k_spectrum<MODE, PREC, SWEEP, M, false>), dim3((unsigned)grid), dim3(threads), lds, st,
                       kp);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const Plan &pl, int prec, const KParams &kp, int64_t grid, int threads,
                       hipStream_t st) {
  if (prec == SR_PREC_F64) {
    if (pl.M == 2) return launch_t<MODE, SR_PREC_F64, SWEEP_DIRECT, 2>(kp, grid, threads, st);
    return launch_t<MODE, SR_PREC_F64, SWEEP_DIRECT, 1>(kp, grid, threads, st);
  }
  if (pl.sweep == SWEEP_REC) {
    if (MODE == SR_MODE_NEAR || pl.M == 2)
      return launch_t<MODE, SR_PREC_SPLIT, SWEEP_REC, 2>(kp, grid, threads, st);
    return launch_t<MODE, SR_PREC_SPLIT, SWEEP_REC, (MODE == SR_MODE_FAR ? 4 : 2)>(kp, grid, threads, st);
  }
  if (MODE == SR_MODE_FAR && pl.M == 4)
    return launch_t<MODE, SR_PREC_SPLIT, SWEEP_DIRECT, (MODE == SR_MODE_FAR ? 4 : 2)>(kp, grid, threads, st);
  if (pl.M == 2) return launch_t<MODE, SR_PREC_SPLIT, SWEEP_DIRECT, 2>(kp, grid, threads, st);
  return launch_t<MODE, SR_PREC_SPLIT, SWEEP_DIRECT, 1>(kp, grid, threads, st);
}

int check_grid(const sr_grid *g) {
  if (!g) return fail(SR_ERR_INVALID, "grid is NULL");
  if (g->mode != SR_MODE_FAR && g->mode != SR_MODE_NEAR)
    return fail(SR_ERR_INVALID, "grid.mode %d is not far (0) or near (1)", g->mode);
  if (g->n_omega == 0 || g->n_theta == 0 || g->n_phi == 0)
    return fail(SR_ERR_INVALID, "grid sizes must be >= 1 (got %u, %u, %u)", g->n_omega, g->n_theta,
                g->n_phi);
  if (!g->omega) return fail(SR_ERR_INVALID, "grid axis pointer is NULL");
  if (g->screen_y && g->mode != SR_MODE_NEAR)
    return fail(SR_ERR_INVALID, "screen_y (a Cartesian screen) is a near-field grid");
  if (!g->screen_y && (!g->sin_phi || !g->cos_phi)) return fail(SR_ERR_INVALID, "grid axis pointer is NULL");
  if (g->mode == SR_MODE_FAR && (!g->sin_theta || !g->cos_theta))
    return fail(SR_ERR_INVALID, "far grid needs sin_theta and cos_theta");
  if (g->mode == SR_MODE_NEAR && !g->radius) return fail(SR_ERR_INVALID, "near grid needs radius");
  if (!isfinite(g->omega_step)) return fail(SR_ERR_INVALID, "omega_step must be finite");
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
  if (!(o->omega_per_lane == 0 || o->omega_per_lane == 1 || o->omega_per_lane == 2 ||
        o->omega_per_lane == 4))
    return fail(SR_ERR_INVALID, "omega_per_lane must be 0, 1, 2 or 4");
  if (!(o->waves_per_block == 0 || o->waves_per_block == 1 || o->waves_per_block == 2 ||
        o->waves_per_block == 4))
    return fail(SR_ERR_INVALID, "waves_per_block must be 0, 1, 2 or 4");
  if (o->slices < 0 || o->slices > kMaxSlices)
    return fail(SR_ERR_INVALID, "slices must be in [0, %d]", kMaxSlices);
  if (o->persistent < 0) return fail(SR_ERR_INVALID, "persistent must be >= 0");
  if (o->tile_order != 0 && o->tile_order != 1) return fail(SR_ERR_INVALID, "tile_order must be 0 or 1");
  if (o->sweep < SR_SWEEP_AUTO || o->sweep > SR_SWEEP_RECURRENCE)
    return fail(SR_ERR_INVALID, "sweep must be SR_SWEEP_AUTO, _DIRECT or _RECURRENCE");
  if (o->sweep == SR_SWEEP_RECURRENCE && (o->precision != SR_PREC_SPLIT || g->omega_step == 0.0))
    return fail(SR_ERR_INVALID, "the recurrence sweep needs SR_PREC_SPLIT and a constant omega_step");
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
  kp.screen_y = g.screen_y;
  kp.omega_step = g.omega_step;
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

// workspace of a batch call: the slices' partial spectra, then the
// direction-independent step data
size_t scratch_bytes(const Plan &pl, int n_out, int64_t n_grid, uint32_t n_snaps) {
  // rounded to 256 bytes: the step records after it are read with 16-byte loads
  return pl.slices > 1 ? ((size_t)pl.slices * n_out * (size_t)n_grid * n_snaps * sizeof(double) + 255) / 256 * 256
                       : 0;
}
size_t pre_bytes(int mode, int64_t n_points) {
  return (size_t)(mode == SR_MODE_FAR ? 6 : 3) * (size_t)n_points * sizeof(double);
}
// the tile costs and launch order (batch calls in the far field with enough
// tiles and work to have a tail worth ordering), 0 otherwise
size_t order_bytes(const Plan &pl, const sr_grid &g, const sr_opts &o, int64_t n_points) {
  const double evals = (double)n_points * g.n_omega * g.n_theta * g.n_phi;
  if (g.mode != SR_MODE_FAR || o.tile_order != 0 || n_points < 2 || pl.blocks < kOrderMinTiles ||
      pl.blocks > kOrderMaxTiles || evals < kOrderMinEvals ||
      evals < kOrderEvalsPerTile2 * (double)pl.blocks * (double)pl.blocks)
    return 0;
  return ((size_t)pl.blocks * (sizeof(float) + sizeof(int32_t)) + 255) / 256 * 256;
}

// want_pre: compute the step data into the workspace (batch calls); the
// single-particle entry points and the no-allocation batch call compute it in
// the stage instead.  need_order: bytes of the tile launch order (0: index
// order), order_bytes() for batch calls.  Workspace layout: the slices'
// partial spectra, the step data, the tile costs and order.
int run(KParams &kp, const Plan &pl, const sr_grid &g, const sr_opts &o, void *ws, size_t ws_bytes,
        hipStream_t st, bool want_pre = false, size_t need_order = 0) {
  const int threads = kWave * pl.waves;
  const int64_t run = (int64_t)threads * pl.M;
  kp.n_chunks = (uint32_t)((g.n_omega + run - 1) / run);
  kp.slices = pl.slices;
  kp.n_tiles = pl.blocks;
  if (pl.blocks >= ((int64_t)1 << 31))
    return fail(SR_ERR_UNSUPPORTED, "%lld tiles in one launch: at most 2^31 - 1", (long long)pl.blocks);
  kp.div_chunks = make_udiv(kp.n_chunks);
  kp.div_dir = make_udiv(kp.n_dir);
  kp.div_theta = make_udiv(kp.n_theta);
  kp.acc_lds = (kp.n_out == 1 && kp.n_snaps == 1) ? 1 : 0;
  // per-particle outputs are disjoint: slices write them in place
  const size_t need = kp.part_stride ? 0 : scratch_bytes(pl, kp.n_out, kp.n_grid, kp.n_snaps);
  const size_t need_pre = want_pre ? pre_bytes(g.mode, kp.n_points) : 0;
  hipError_t e;
  if (need + need_pre + need_order && (!ws || ws_bytes < need + need_pre + need_order))
    return fail(SR_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need + need_pre + need_order,
                ws_bytes);
  kp.tile_order = nullptr;
  if (need_order) {
    float *cost = (float *)((char *)ws + need + need_pre);
    int32_t *order = (int32_t *)(cost + pl.blocks);
    const int64_t pairs = (int64_t)kp.n_dir * kp.n_chunks;
    hipLaunchKernelGGL(k_tile_cost, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, st, kp, run, cost);
    hipLaunchKernelGGL(k_tile_rank, dim3((unsigned)((pl.blocks + 255) / 256)), dim3(256), 0, st,
                       (const float *)cost, (int64_t)pl.blocks, order);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SR_ERR_HIP, "tile order launch: %s", hipGetErrorString(e));
    kp.tile_order = order;
  }
  if (need_pre) {
    kp.pre = (const double *)((char *)ws + need);
    const int64_t n = kp.n_points;
    if (g.mode == SR_MODE_FAR)
      hipLaunchKernelGGL(k_step_prep<SR_MODE_FAR>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kp.ux,
                         kp.uy, kp.uz, n, kp.dt, (double *)kp.pre);
    else
      hipLaunchKernelGGL(k_step_prep<SR_MODE_NEAR>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kp.ux,
                         kp.uy, kp.uz, n, kp.dt, (double *)kp.pre);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SR_ERR_HIP, "step data launch: %s", hipGetErrorString(e));
  }
  if (need) {
    kp.scratch = (double *)ws;
    const int64_t n = (int64_t)(need / sizeof(double));
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kp.scratch, n);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SR_ERR_HIP, "workspace clear: %s", hipGetErrorString(e));
  } else {
    kp.scratch = nullptr;
  }
  int64_t grid = pl.blocks;
  if (o.persistent > 0) grid = std::min<int64_t>(grid, o.persistent);
  if (grid <= 0) return SR_OK;
  e = (g.mode == SR_MODE_FAR) ? launch_mode<SR_MODE_FAR>(pl, o.precision, kp, grid, threads, st)
                              : launch_mode<SR_MODE_NEAR>(pl, o.precision, kp, grid, threads, st);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "tile kernel launch: %s", hipGetErrorString(e));
  if (need) {
    const int64_t total = (int64_t)kp.n_out * kp.n_grid * kp.n_snaps;
    hipLaunchKernelGGL(k_reduce_slices, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, kp);
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
  pl.slices = 1;  // one particle: no slices, no workspace
  slice_particles(kp, 1, 1, nullptr);
  return run(kp, pl, g, o, nullptr, 0, (hipStream_t)stream);
}

}  // namespace

namespace sr {

int step_prep(const sr_tracks &t, int32_t mode, double *pre, hipStream_t st) {
  const int64_t n = t.n_points;
  if (mode == SR_MODE_FAR)
    hipLaunchKernelGGL(k_step_prep<SR_MODE_FAR>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t.ux, t.uy,
                       t.uz, n, t.dt, pre);
  else
    hipLaunchKernelGGL(k_step_prep<SR_MODE_NEAR>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t.ux, t.uy,
                       t.uz, n, t.dt, pre);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? SR_OK : fail(SR_ERR_HIP, "step data launch: %s", hipGetErrorString(e));
}

namespace {
// samples of particles [p_begin, p_end): exact with the host offsets, else
// the batch's mean
int64_t chunk_points(const sr_tracks &t, int64_t p_begin, int64_t p_end) {
  if (t.offsets_host) return t.offsets_host[p_end] - t.offsets_host[p_begin];
  return t.n_particles ? t.n_points * (p_end - p_begin) / t.n_particles : 0;
}
sr_opts fields_opts(int32_t precision) {
  sr_opts o;
  memset(&o, 0, sizeof(o));
  o.comp = kCompField;
  o.precision = precision;
  // every particle's field has its own output, so particle slices cost no
  // reduction: split fields take the finest tiles (the adjoint workload's
  // 256 x 4000 steps on 128 x 32 x 32: 41.4 ms at 32 slices against 42.7 at
  // the library plan's 8, heaviest first); FP64 keeps the plan (90.6 ms
  // against 93.1 at 32)
  if (precision == SR_PREC_SPLIT) o.slices = kMaxSlices;
  return o;
}
}  // namespace

size_t fields_order_space(const sr_grid &g, int32_t precision, int64_t n_particles) {
  const Plan pl = make_plan(g, fields_opts(precision), std::max<int64_t>(n_particles, 1), 1);
  if (pl.blocks > kOrderMaxTiles) return 0;
  return ((size_t)pl.blocks * (sizeof(float) + sizeof(int32_t)) + 255) / 256 * 256;
}

int spectrum_fields(const sr_tracks &t, const sr_grid &g, int32_t precision, int64_t p_begin, int64_t p_end,
                    const double *pre, double *F, hipStream_t st, void *order_ws, size_t order_ws_bytes) {
  KParams kp;
  memset(&kp, 0, sizeof(kp));
  kp.offsets = t.offsets;
  kp.x = t.x; kp.y = t.y; kp.z = t.z; kp.ux = t.ux; kp.uy = t.uy; kp.uz = t.uz;
  kp.weight = t.weight;
  kp.it_start = t.it_start;
  kp.it_end = t.it_end;
  kp.it_snaps = t.it_snaps;
  kp.n_snaps = 1;
  kp.snaps_stride = t.snaps_stride;
  kp.dt = t.dt;
  kp.n_points = t.n_points;
  kp.pre = pre;
  const sr_opts o = fields_opts(precision);
  const int64_t np = p_end - p_begin;
  const int64_t n_grid = (int64_t)g.n_omega * g.n_theta * g.n_phi;
  double *outs[kMaxOut];
  for (int c = 0; c < 6; c++) outs[c] = F + (int64_t)c * np * n_grid;
  fill_grid(kp, g, o, 6, outs);
  const Plan pl = make_plan(g, o, np, 1);
  for (int q = 0; q <= pl.slices; q++) kp.slice_lo[q] = p_begin + np * q / pl.slices;
  kp.part_base = p_begin;
  kp.part_stride = n_grid;
  // heaviest tiles first where the launch is large enough and the caller
  // gave room for the order
  size_t ob = order_ws ? order_bytes(pl, g, o, chunk_points(t, p_begin, p_end)) : 0;
  if (ob > order_ws_bytes) ob = 0;
  return run(kp, pl, g, o, order_ws, order_ws_bytes, st, false, ob);
}

}  // namespace sr

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int sr_abi_version(void) { return SR_ABI_VERSION; }

const char *sr_last_error(void) { return sr::last_error(); }

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
  if (tracks->n_particles == 0) return SR_OK;
  const Plan pl = make_plan(*grid, *opts, tracks->n_particles, tracks->n_snaps, tracks->n_points);
  *bytes = scratch_bytes(pl, outputs_for(grid->mode, opts->comp),
                         (int64_t)grid->n_omega * grid->n_theta * grid->n_phi, tracks->n_snaps) +
           pre_bytes(grid->mode, tracks->n_points) + order_bytes(pl, *grid, *opts, tracks->n_points);
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
  kp.n_points = t->n_points;
  fill_grid(kp, *g, *o, n_out, outs);
  const Plan pl = make_plan(*g, *o, t->n_particles, t->n_snaps, t->n_points);
  slice_particles(kp, pl.slices, t->n_particles, t->offsets_host);
  return run(kp, pl, *g, *o, workspace, workspace_bytes, (hipStream_t)stream, true,
             order_bytes(pl, *g, *o, t->n_points));
}

int sr_plan(const sr_tracks *t, const sr_grid *g, const sr_opts *o, sr_plan_info *info) {
  int rc;
  if (!info) return fail(SR_ERR_INVALID, "info is NULL");
  memset(info, 0, sizeof(*info));
  if ((rc = check_tracks(t)) || (rc = check_grid(g)) || (rc = check_opts(g, o))) return rc;
  const Plan pl = make_plan(*g, *o, std::max<int64_t>(t->n_particles, 1), t->n_snaps, t->n_points);
  info->sweep = pl.sweep == SWEEP_REC ? SR_SWEEP_RECURRENCE : SR_SWEEP_DIRECT;
  info->omega_per_lane = pl.M;
  info->waves_per_block = pl.waves;
  info->slices = pl.slices;
  info->tiles = pl.blocks;
  info->ordered = order_bytes(pl, *g, *o, t->n_points) > 0 ? 1 : 0;
  return SR_OK;
}

namespace {
// the candidate's plan fields over the base request's output and precision
sr_opts candidate_opts(const sr_opts &base, const sr_opts &cand) {
  sr_opts o = cand;
  o.comp = base.comp;
  o.precision = base.precision;
  o.form_factor = base.form_factor;
  return o;
}
}  // namespace

int sr_autotune_workspace_bytes(const sr_tracks *t, const sr_grid *g, const sr_opts *base,
                                const sr_opts *candidates, int n_candidates, size_t *bytes) {
  if (!bytes) return fail(SR_ERR_INVALID, "bytes is NULL");
  *bytes = 0;
  if (!base || !candidates || n_candidates <= 0) return fail(SR_ERR_INVALID, "no candidates");
  for (int i = 0; i < n_candidates; i++) {
    const sr_opts o = candidate_opts(*base, candidates[i]);
    size_t b = 0;
    if (sr_workspace_bytes(t, g, &o, &b) == SR_OK) *bytes = std::max(*bytes, b);
  }
  return SR_OK;
}

int sr_autotune(const sr_tracks *t, const sr_grid *g, const sr_opts *base, const sr_opts *candidates,
                int n_candidates, double *const *outs, void *workspace, size_t workspace_bytes, int warmup,
                int reps, void *stream, float *ms, int *best) {
  int rc;
  if (!base || !candidates || n_candidates <= 0 || !ms || !best)
    return fail(SR_ERR_INVALID, "sr_autotune needs a base request, candidates, ms[] and best");
  if ((rc = check_tracks(t)) || (rc = check_grid(g)) || (rc = check_opts(g, base))) return rc;
  if (reps < 1 || warmup < 0) return fail(SR_ERR_INVALID, "reps must be >= 1 and warmup >= 0");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return fail(SR_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(SR_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  *best = -1;
  float best_ms = INFINITY;
  rc = SR_OK;
  for (int i = 0; i < n_candidates && rc == SR_OK; i++) {
    ms[i] = INFINITY;
    const sr_opts o = candidate_opts(*base, candidates[i]);
    size_t need = 0;
    if (check_opts(g, &o) != SR_OK || sr_workspace_bytes(t, g, &o, &need) != SR_OK || need > workspace_bytes)
      continue;   // a plan this request cannot run (or cannot fit) is not a candidate
    // untimed runs (code objects, caches), then reps timed ones
    for (int r = 0; r < warmup && rc == SR_OK; r++)
      rc = sr_spectrum_batch(t, g, &o, outs, workspace, workspace_bytes, stream);
    if (rc != SR_OK) break;
    e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && rc == SR_OK && e == hipSuccess; r++)
      rc = sr_spectrum_batch(t, g, &o, outs, workspace, workspace_bytes, stream);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float elapsed = 0.f;
    if (e == hipSuccess && rc == SR_OK) e = hipEventElapsedTime(&elapsed, e0, e1);
    if (e != hipSuccess) {
      rc = fail(SR_ERR_HIP, "timing candidate %d: %s", i, hipGetErrorString(e));
      break;
    }
    if (rc != SR_OK) break;
    ms[i] = elapsed / (float)reps;
    if (ms[i] < best_ms) {
      best_ms = ms[i];
      *best = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == SR_OK && *best < 0) rc = fail(SR_ERR_INVALID, "no candidate plan can run this request");
  return rc;
}

int sr_spectrum_batch_noalloc(const sr_tracks *t, const sr_grid *g, const sr_opts *o, double *const *outs,
                              void *stream) {
  int rc;
  if ((rc = check_tracks(t)) || (rc = check_grid(g)) || (rc = check_opts(g, o))) return rc;
  if (t->n_particles == 0) return SR_OK;
  const int n_out = outputs_for(g->mode, o->comp);
  if (!outs) return fail(SR_ERR_INVALID, "outs is NULL");
  for (int c = 0; c < n_out; c++)
    if (!outs[c]) return fail(SR_ERR_INVALID, "output %d is NULL", c);
  void *scr = nullptr;
  if ((rc = noalloc_scratch(&scr))) return rc;
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
  kp.n_points = t->n_points;
  fill_grid(kp, *g, *o, n_out, outs);
  Plan pl = make_plan(*g, *o, t->n_particles, t->n_snaps, t->n_points);
  // the same plan as sr_spectrum_batch; a tuned slice count is halved until
  // the slices' partial spectra and the tile order fit the scratch array
  auto need = [&](const Plan &p) {
    return scratch_bytes(p, n_out, kp.n_grid, t->n_snaps) + order_bytes(p, *g, *o, t->n_points);
  };
  while (o->slices == 0 && pl.slices > 1 && need(pl) > kNoAllocScratchBytes) {
    pl.blocks /= pl.slices;
    pl.slices /= 2;
    pl.blocks *= pl.slices;
  }
  if (need(pl) > kNoAllocScratchBytes)
    return fail(SR_ERR_WORKSPACE,
                "this request needs %zu bytes of scratch (%d particle slices); the no-allocation entry "
                "point holds %zu: ask for fewer slices (opts.slices) or use sr_spectrum_batch",
                need(pl), pl.slices, kNoAllocScratchBytes);
  slice_particles(kp, pl.slices, t->n_particles, t->offsets_host);
  return run(kp, pl, *g, *o, scr, kNoAllocScratchBytes, (hipStream_t)stream, false,
             order_bytes(pl, *g, *o, t->n_points));
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
