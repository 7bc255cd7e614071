// common.h -- constants, device helpers and host error reporting shared by the
// spectrum and adjoint kernels of synchrad_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <math.h>

#include "synchrad_amd.h"

namespace sr {

constexpr int kWave = 64;
constexpr int kMaxT = 256;        // threads per block = steps per LDS tile
constexpr int kMaxSlices = 32;
constexpr int kMaxOut = 6;

constexpr double kPi = 3.141592653589793;
constexpr double kInv2Pi = 0.15915494309189535;
constexpr double k2PiHi = 6.283185307179586;
constexpr double k2PiLo = 2.4492935982947064e-16;

// Records a printf-style message for sr_last_error() and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Internal output kind of the tile kernel: the complex far field
// F_p(g) = sum_s A_s exp(i phi_s) of each particle (no weight, no dt), as
// x re, x im, y re, y im, z re, z im.
constexpr int kCompField = 100;

// The direction-independent step data of every sample into pre (far
// [a_x, a_y, a_z, bbar_x, bbar_y, bbar_z][n_points], near [b_x, b_y, b_z]
// [n_points]), as the tile kernel's stage would compute it (spectrum.hip).
int step_prep(const sr_tracks &t, int32_t mode, double *pre, hipStream_t st);

// Adds the fields of particles [p_begin, p_end) (one snapshot record each) into
// F[6][p_end - p_begin][n_grid] with the tile kernel's own plan for `precision`
// (spectrum.hip); pre is step_prep's output or nullptr.  F must be cleared by
// the caller; stream-ordered.  order_ws (fields_order_space bytes for the
// largest chunk, or nullptr) lets large launches run their tiles heaviest first.
int spectrum_fields(const sr_tracks &t, const sr_grid &g, int32_t precision, int64_t p_begin, int64_t p_end,
                    const double *pre, double *F, hipStream_t st, void *order_ws = nullptr,
                    size_t order_ws_bytes = 0);
// workspace for the tile order of a fields launch of n_particles particles
size_t fields_order_space(const sr_grid &g, int32_t precision, int64_t n_particles);

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double unid(double v) {
  return __longlong_as_double(uni64(__double_as_longlong(v)));
}
// The value unchanged, but opaque to the optimiser from this point: addresses
// formed from it are formed here, not hoisted out of the loops around (where
// the compiler would keep one 64-bit per-lane pointer per output and spill
// them across the sweep).  An empty asm statement: it emits no instruction.
template <typename V>
__device__ __forceinline__ V late(V v) {
  asm volatile("" : "+v"(v));
  return v;
}
// The same for a wave-uniform value, kept in a scalar register: a flag read
// here is tested here by a scalar branch, not carried through the loops as a
// per-lane copy.
template <typename V>
__device__ __forceinline__ V late_s(V v) {
  asm volatile("" : "+s"(v));
  return v;
}

// Division of n < 2^31 by a divisor fixed for a launch, as a multiply and a
// shift (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994, Thm 4.2 with N = 31): l = ceil(log2 d),
// m = ceil(2^(31+l) / d) < 2^32 gives floor(n / d) = (n m) >> (31 + l).  With
// n and the divisor wave-uniform this runs on the scalar unit; the compiler's
// own division would form its reciprocal in vector registers and hold it
// across the tile loop.
struct UDiv {
  uint32_t m, sh;
};
inline UDiv make_udiv(uint32_t d) {
  if (d == 0) return UDiv{0u, 31u};
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  const uint64_t m = ((1ull << (31 + l)) + d - 1) / d;
  return UDiv{(uint32_t)m, 31u + l};
}
__host__ __device__ __forceinline__ uint32_t udiv(uint32_t n, UDiv v) {
  return (uint32_t)(((uint64_t)n * v.m) >> v.sh);
}

// FP64 sin and cos of a phase.  The quadrant k = rint(2 ph / pi) is removed
// with pi/2 split in two (hi = fl(pi/2), lo = pi/2 - hi; the FMA forms
// ph - k hi exactly before rounding), leaving |r| <= pi/4 (+ rounding) for
// |k| < 2^26.  sin r and cos r are their Taylor series to r^17 and r^16 (the
// first omitted terms are below 1e-19 there), evaluated by Horner in FMA; the
// quadrant then picks and signs them.  Accurate to a few units in the last
// place, which the FP64 sums' own rounding dominates.
#ifndef SR_F64_SINCOS
#define SR_F64_SINCOS 1
#endif
constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPiHalfHi = 1.5707963267948966;
constexpr double kPiHalfLo = 6.123233995736766e-17;

__device__ __forceinline__ void sincos_reduced(double ph, double *s, double *c) {
#if SR_F64_SINCOS
  const double k = rint(ph * kTwoOverPi);
  double r = fma(-k, kPiHalfHi, ph);
  r = fma(-k, kPiHalfLo, r);
  const double r2 = r * r;
  // 1/n! for n = 17, 15, ..., 3 and 16, 14, ..., 2
  double ps = 2.8114572543455206e-15;
  ps = fma(ps, r2, -7.647163731819816e-13);
  ps = fma(ps, r2, 1.6059043836821613e-10);
  ps = fma(ps, r2, -2.505210838544172e-08);
  ps = fma(ps, r2, 2.7557319223985893e-06);
  ps = fma(ps, r2, -0.0001984126984126984);
  ps = fma(ps, r2, 0.008333333333333333);
  ps = fma(ps, r2, -0.16666666666666666);
  const double sr = fma(ps * r2, r, r);
  double pc = 4.779477332387385e-14;
  pc = fma(pc, r2, -1.1470745597729725e-11);
  pc = fma(pc, r2, 2.08767569878681e-09);
  pc = fma(pc, r2, -2.755731922398589e-07);
  pc = fma(pc, r2, 2.48015873015873e-05);
  pc = fma(pc, r2, -0.001388888888888889);
  pc = fma(pc, r2, 0.041666666666666664);
  pc = fma(pc, r2, -0.5);
  const double cr = fma(pc, r2, 1.0);
  const int q = (int)(long long)k & 3;
  const double sq = (q & 1) ? cr : sr;
  const double cq = (q & 1) ? sr : cr;
  *s = (q & 2) ? -sq : sq;
  *c = ((q + 1) & 2) ? -cq : cq;
#else
  const double k = rint(ph * kInv2Pi);
  double r = fma(-k, k2PiHi, ph);
  r = fma(-k, k2PiLo, r);
  sincos(r, s, c);
#endif
}

// FP64 sin and cos of a phase from a table of sin/cos(i pi/128), i = 0..255
// (kSinCosTab entries, held in LDS by the caller).  k = rint(128 ph / pi) comes
// from one FMA with 1.5 * 2^52 added (the sum's low mantissa bits are then k,
// so k mod 256 is read off its low word, for |ph| < kTabMaxPhase; when `wide`
// (uniform: the caller's bound on |ph| reaches kTabMaxPhase / 2) the phase is
// first brought into [-pi, pi] with 2 pi split in two); the remainder
// r = ph - k pi/128 is formed with pi/128 split in two as above, |r| <= pi/256;
// sin r to r^5 and cos r to r^6 (first omitted terms below 1e-17 there); then
// sin ph = S_k cos r + C_k sin r and cos ph = C_k cos r - S_k sin r.  Against
// 40-digit values on 2.5e4 phases up to 1e9 rad the largest error is 1.5e-16.
#ifndef SR_F64_TABLE
#define SR_F64_TABLE 1
#endif
constexpr int kSinCosTab = 256;
constexpr double kTabPerRad = 40.74366543152521;      // 128 / pi
constexpr double kTabStepHi = 0.02454369260617026;    // fl(pi / 128)
constexpr double kTabStepLo = 9.567553118338697e-19;  // pi / 128 - kTabStepHi
constexpr double kRintShift = 6755399441055744.0;     // 1.5 * 2^52
constexpr double kTabMaxPhase = 17592186044416.0;     // 2^44 (the shift needs |128 ph/pi| < 2^51)

__device__ __forceinline__ void sincos_table(double ph, const double2 *tab, bool wide, double *s, double *c) {
  if (wide) {
    const double n = rint(ph * kInv2Pi);
    ph = fma(-n, k2PiLo, fma(-n, k2PiHi, ph));
  }
  const double y = fma(ph, kTabPerRad, kRintShift);
  const double k = y - kRintShift;
  const double2 t = tab[(uint32_t)__double_as_longlong(y) & (kSinCosTab - 1)];
  double r = fma(-k, kTabStepHi, ph);
  r = fma(-k, kTabStepLo, r);
  const double r2 = r * r;
  const double sr = fma(r * r2, fma(r2, 0.008333333333333333, -0.16666666666666666), r);
  const double cr = fma(fma(fma(r2, -0.001388888888888889, 0.041666666666666664), r2, -0.5), r2, 1.0);
  *s = fma(t.x, cr, t.y * sr);
  *c = fma(t.y, cr, -(t.x * sr));
}

// sin/cos of 2*pi*rv from an FP64 count of revolutions: the fraction is taken
// exactly in FP64, then the FP32 hardware units (input in revolutions) run.
// component q (0..3, a compile-time constant after unrolling) of a float4
__device__ __forceinline__ float lane4(const float4 &v, int q) {
  return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
}

__device__ __forceinline__ void sincos_rev(double rv, float *s, float *c) {
  const float f = (float)__builtin_amdgcn_fract(rv);
  *s = __builtin_amdgcn_sinf(f);
  *c = __builtin_amdgcn_cosf(f);
}

// 1/sqrt(1 + u.u) as the reference's rsqrt(1 + dot(u,u)), kernel_farfield.cl:77
__device__ __forceinline__ double gamma_inv(double ux, double uy, double uz) {
  return 1.0 / sqrt(1.0 + (ux * ux + uy * uy + uz * uz));
}

// a = (b_{i+1} - b_i)/dt and bbar = (b_{i+1} + b_i)/2 of sample i's interval,
// b = u rsqrt(1 + u.u), in the reference's expression order
// (kernel_farfield.cl:74-83)
__device__ __forceinline__ void step_am(const double *ux, const double *uy, const double *uz, int64_t i,
                                        double dtInv, double *ax, double *ay, double *az, double *mx,
                                        double *my, double *mz) {
  const double u0x = ux[i], u0y = uy[i], u0z = uz[i];
  const double u1x = ux[i + 1], u1y = uy[i + 1], u1z = uz[i + 1];
  const double g0 = gamma_inv(u0x, u0y, u0z), g1 = gamma_inv(u1x, u1y, u1z);
  const double b0x = u0x * g0, b0y = u0y * g0, b0z = u0z * g0;
  const double b1x = u1x * g1, b1y = u1y * g1, b1z = u1z * g1;
  *ax = (b1x - b0x) * dtInv; *ay = (b1y - b0y) * dtInv; *az = (b1z - b0z) * dtInv;
  *mx = 0.5 * (b1x + b0x); *my = 0.5 * (b1y + b0y); *mz = 0.5 * (b1z + b0z);
}

struct Particle {
  int64_t off, n;
  double w;
  uint32_t its, ite;
  const uint32_t *snaps;
};

template <class Params>
__device__ __forceinline__ Particle load_particle(const Params &p, int64_t i) {
  Particle r;
  if (p.offsets) {
    r.off = uni64(p.offsets[i]);
    r.n = uni64(p.offsets[i + 1]) - r.off;
    r.w = unid(p.weight[i]);
    r.its = (uint32_t)uni((int)p.it_start[i]);
    r.ite = (uint32_t)uni((int)p.it_end[i]);
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
    const int64_t r = (int64_t)(uint32_t)uni((int)pt.snaps[slot]) - 2 - (int64_t)pt.its;
    if (r <= last || r > (int64_t)pt.ite - 2) return;
    next = r;
  }
  __device__ void init(const Particle &pt, uint32_t n_snaps) {
    slot = 0;
    while (slot < n_snaps && !(pt.its < (uint32_t)uni((int)pt.snaps[slot]))) slot++;
    advance(pt, n_snaps, -1);
  }
};

}  // namespace sr
