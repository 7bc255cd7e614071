/*
 * oracle.c -- CPU FP64 restatement of synchrad's spectral radiation integral.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels of
 * synchrad_amd are compared against.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * It restates, one work item per spectral point and one call per particle, the
 * arithmetic of the reference's OpenCL kernels:
 *   far  total                 synchrad/kernel_farfield.cl:6-109
 *   far  cartesian_comps       synchrad/kernel_farfield.cl:111-225
 *   far  cartesian_complex     synchrad/kernel_farfield.cl:227-344
 *   far  spheric_comps         synchrad/kernel_farfield.cl:346-469
 *   far  spheric_complex       synchrad/kernel_farfield.cl:471-595
 *   near total                 synchrad/kernel_nearfield.cl:5-104
 *   near cartesian_comps       synchrad/kernel_nearfield.cl:106-214
 *   near cartesian_complex     synchrad/kernel_nearfield.cl:216-326
 * with the argument order of those kernels (far: kernel_farfield.cl:6-28,
 * near: kernel_nearfield.cl:5-27).
 *
 * Conventions fixed here (the OpenCL source leaves them to the runtime):
 *   - dot(a,b) is evaluated left to right, (ax*bx + ay*by) + az*bz;
 *   - no floating-point contraction (built with -ffp-contract=off);
 *   - rsqrt(v) is 1.0/sqrt(v) with IEEE sqrt and division;
 *   - sin/cos are the C library's double routines;
 *   - after the last snapshot the kernel reads itSnaps[nSnaps] out of bounds
 *     (kernel_farfield.cl:100); here that read is taken as "no record".
 * Inputs the reference would hang or read out of bounds on (itEnd == 0 or
 * nSteps == 0, kernel_farfield.cl:59,63) are rejected with a negative return.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
  OR_TOTAL = 0,
  OR_CARTESIAN = 1,
  OR_CARTESIAN_COMPLEX = 2,
  OR_SPHERIC = 3,
  OR_SPHERIC_COMPLEX = 4
};

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 add3(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl3(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }

int oracle_abi_version(void) { return 1; }

/* First snapshot index with itStart < itSnaps[i]  (kernel_farfield.cl:53-57). */
static uint32_t first_snap(uint32_t itStart, uint32_t nSnaps, const uint32_t *itSnaps) {
  uint32_t i;
  for (i = 0; i < nSnaps; i++)
    if (itStart < itSnaps[i]) break;
  return i;
}

static int n_components(int comp) {
  switch (comp) {
    case OR_TOTAL: return 1;
    case OR_CARTESIAN: case OR_SPHERIC: return 3;
    case OR_CARTESIAN_COMPLEX: case OR_SPHERIC_COMPLEX: return 6;
    default: return -1;
  }
}

/* Snapshot record (kernel_farfield.cl:100-106, 208-222, 324-341). */
static void record(int comp, double *const *spec, size_t idx, double wpdt2, double wpdt,
                   v3 re, v3 im) {
  switch (comp) {
    case OR_TOTAL:
      spec[0][idx] += wpdt2 * (dot3(re, re) + dot3(im, im));
      break;
    case OR_CARTESIAN:
    case OR_SPHERIC:
      spec[0][idx] += wpdt2 * (re.x * re.x + im.x * im.x);
      spec[1][idx] += wpdt2 * (re.y * re.y + im.y * im.y);
      spec[2][idx] += wpdt2 * (re.z * re.z + im.z * im.z);
      break;
    default: /* complex: += sqrt(w)*dt*(Re, Im) per component */
      spec[0][idx] += wpdt * re.x;
      spec[1][idx] += wpdt * im.x;
      spec[2][idx] += wpdt * re.y;
      spec[3][idx] += wpdt * im.y;
      spec[4][idx] += wpdt * re.z;
      spec[5][idx] += wpdt * im.z;
      break;
  }
}

/*
 * Far field, one particle.  Argument order follows kernel_farfield.cl:6-28;
 * `comp` selects the kernel, `spectra` holds 1, 3 or 6 output arrays laid out
 * (nSnaps, nPhi, nTheta, nOmega), and `formFactor` (nOmega) is read only by
 * the cartesian_complex kernel (kernel_farfield.cl:271, 324-325).
 */
int oracle_far(int comp, double *const *spectra,
               const double *x, const double *y, const double *z,
               const double *ux, const double *uy, const double *uz,
               double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
               const double *omega, const double *sinTheta, const double *cosTheta,
               const double *sinPhi, const double *cosPhi,
               uint32_t nOmega, uint32_t nTheta, uint32_t nPhi,
               double dt, uint32_t nSnaps, const uint32_t *itSnaps,
               const double *formFactor, int nthreads) {
  if (n_components(comp) < 0) return -1;
  if (itEnd == 0 || nSteps == 0) return -2;
  if (comp == OR_CARTESIAN_COMPLEX && formFactor == NULL) return -3;
  const int64_t nTotal = (int64_t)nTheta * nPhi * nOmega;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

#pragma omp parallel for schedule(static)
  for (int64_t gti = 0; gti < nTotal; gti++) {
    /* kernel_farfield.cl:35-42 */
    const uint32_t iPhi = (uint32_t)(gti / ((int64_t)nOmega * nTheta));
    const uint32_t iTheta = (uint32_t)((gti - (int64_t)iPhi * nOmega * nTheta) / nOmega);
    const uint32_t iOmega = (uint32_t)(gti - (int64_t)iPhi * nOmega * nTheta - (int64_t)iTheta * nOmega);
    const double omegaLocal = omega[iOmega];
    const v3 nVec = mk(sinTheta[iTheta] * cosPhi[iPhi], sinTheta[iTheta] * sinPhi[iPhi],
                       cosTheta[iTheta]);
    /* spheric basis, kernel_farfield.cl:384-387 */
    const v3 thVec = mk(cosTheta[iTheta] * cosPhi[iPhi], cosTheta[iTheta] * sinPhi[iPhi],
                        -sinTheta[iTheta]);
    const v3 phVec = mk(-sinPhi[iPhi], cosPhi[iPhi], 0.0);
    const double ffLocal = (comp == OR_CARTESIAN_COMPLEX) ? formFactor[iOmega] : 1.0;

    const double dtInv = 1.0 / dt;
    const double wpdt2 = wp * dt * dt;
    const double wpdt = sqrt(wp) * dt;
    double phasePrev = 0.0;
    v3 re = mk(0, 0, 0), im = mk(0, 0, 0);

    uint32_t iSnap = first_snap(itStart, nSnaps, itSnaps);

    for (uint32_t it = 0; it < itEnd - 1; it++) {
      const uint32_t it_glob = itStart + it;
      if (it < nSteps - 1) {
        /* kernel_farfield.cl:65-72 */
        const double time = (double)it_glob * dt;
        const v3 xl = mk(x[it], y[it], z[it]);
        const double phase = omegaLocal * (time - dot3(xl, nVec));
        const double dPhase = fabs(phase - phasePrev);
        phasePrev = phase;
        if (dPhase < M_PI) {
          /* kernel_farfield.cl:74-96 */
          v3 u = mk(ux[it], uy[it], uz[it]);
          v3 un = mk(ux[it + 1], uy[it + 1], uz[it + 1]);
          double gammaInv = 1.0 / sqrt(1.0 + dot3(u, u));
          u = scl3(gammaInv, u);
          gammaInv = 1.0 / sqrt(1.0 + dot3(un, un));
          un = scl3(gammaInv, un);
          const v3 a = scl3(dtInv, sub3(un, u));
          const v3 ub = scl3(0.5, add3(un, u));
          double c1 = dot3(a, nVec);
          double c2 = 1.0 - dot3(ub, nVec);
          c2 = 1.0 / c2;
          c1 = c1 * c2 * c2;
          const double sinPhase = sin(phase);
          const double cosPhase = cos(phase);
          v3 amp = sub3(scl3(c1, sub3(nVec, ub)), scl3(c2, a));
          if (comp == OR_SPHERIC || comp == OR_SPHERIC_COMPLEX) {
            /* kernel_farfield.cl:442-445 */
            amp = mk(dot3(nVec, amp), dot3(thVec, amp), dot3(phVec, amp));
          }
          if (comp == OR_CARTESIAN_COMPLEX) {
            /* (amplitude*cos)*FormFactor, kernel_farfield.cl:324-325 */
            re = add3(re, scl3(ffLocal, scl3(cosPhase, amp)));
            im = add3(im, scl3(ffLocal, scl3(sinPhase, amp)));
          } else {
            re = add3(re, scl3(cosPhase, amp));
            im = add3(im, scl3(sinPhase, amp));
          }
        }
      }
      if (iSnap < nSnaps && it_glob + 2 == itSnaps[iSnap]) {
        record(comp, spectra, (size_t)(gti + nTotal * iSnap), wpdt2, wpdt, re, im);
        iSnap += 1;
      }
    }
  }
  return 0;
}

/*
 * Near field, one particle, polar screen at distance L.  Argument order
 * follows kernel_nearfield.cl:5-27.  The reference's near cartesian_complex
 * kernel declares a FormFactor argument but never reads it
 * (kernel_nearfield.cl:216-326); it is ignored here as well.
 */
int oracle_near(int comp, double *const *spectra,
                const double *x, const double *y, const double *z,
                const double *ux, const double *uy, const double *uz,
                double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                const double *omega, const double *radius,
                const double *sinPhi, const double *cosPhi, double distanceToScreen,
                uint32_t nOmega, uint32_t nRadius, uint32_t nPhi,
                double dt, uint32_t nSnaps, const uint32_t *itSnaps, int nthreads) {
  if (!(comp == OR_TOTAL || comp == OR_CARTESIAN || comp == OR_CARTESIAN_COMPLEX)) return -1;
  if (itEnd == 0 || nSteps == 0) return -2;
  const int64_t nTotal = (int64_t)nRadius * nPhi * nOmega;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

#pragma omp parallel for schedule(static)
  for (int64_t gti = 0; gti < nTotal; gti++) {
    /* kernel_nearfield.cl:34-42 */
    const uint32_t iPhi = (uint32_t)(gti / ((int64_t)nOmega * nRadius));
    const uint32_t iRadius = (uint32_t)((gti - (int64_t)iPhi * nOmega * nRadius) / nOmega);
    const uint32_t iOmega = (uint32_t)(gti - (int64_t)iPhi * nOmega * nRadius - (int64_t)iRadius * nOmega);
    const double omegaLocal = omega[iOmega];
    const v3 scr = mk(radius[iRadius] * cosPhi[iPhi], radius[iRadius] * sinPhi[iPhi],
                      distanceToScreen);

    const double wpdt2 = wp * dt * dt;
    const double wpdt = sqrt(wp) * dt;
    double phasePrev = 0.0;
    v3 re = mk(0, 0, 0), im = mk(0, 0, 0);
    uint32_t iSnap = first_snap(itStart, nSnaps, itSnaps);

    for (uint32_t it = 0; it < itEnd - 1; it++) {
      const uint32_t it_glob = itStart + it;
      if (it < nSteps - 1) {
        /* kernel_nearfield.cl:64-74 */
        const double time = (double)it_glob * dt;
        const v3 xl = mk(x[it], y[it], z[it]);
        const v3 rVec = sub3(scr, xl);
        const double rLocal = sqrt(dot3(rVec, rVec));
        const double phase = omegaLocal * (time + rLocal);
        const double dPhase = fabs(phase - phasePrev);
        phasePrev = phase;
        if (dPhase < M_PI) {
          /* kernel_nearfield.cl:76-91 */
          const double rInv = 1.0 / rLocal;
          const v3 nVec = scl3(rInv, rVec);
          v3 u = mk(ux[it], uy[it], uz[it]);
          const double gammaInv = 1.0 / sqrt(1.0 + dot3(u, u));
          u = scl3(gammaInv, u);
          const double sinPhase = sin(phase);
          const double cosPhase = cos(phase);
          const v3 c1 = scl3(omegaLocal * rInv, sub3(u, nVec));
          const v3 c2 = scl3(rInv * rInv, nVec);
          re = add3(re, add3(scl3(-sinPhase, c1), scl3(cosPhase, c2)));
          im = add3(im, add3(scl3(cosPhase, c1), scl3(sinPhase, c2)));
        }
      }
      if (iSnap < nSnaps && it_glob + 2 == itSnaps[iSnap]) {
        record(comp, spectra, (size_t)(gti + nTotal * iSnap), wpdt2, wpdt, re, im);
        iSnap += 1;
      }
    }
  }
  return 0;
}
