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
#include <stdlib.h>
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
/* one screen point scr and frequency omegaLocal of the near field: the body of
 * kernel_nearfield.cl:44-101 for spectral index gti */
static void near_point(int comp, double *const *spectra, int64_t gti, int64_t nTotal, v3 scr, double omegaLocal,
                       const double *x, const double *y, const double *z,
                       const double *ux, const double *uy, const double *uz,
                       double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                       double dt, uint32_t nSnaps, const uint32_t *itSnaps) {
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
    const v3 scr = mk(radius[iRadius] * cosPhi[iPhi], radius[iRadius] * sinPhi[iPhi],
                      distanceToScreen);
    near_point(comp, spectra, gti, nTotal, scr, omega[iOmega], x, y, z, ux, uy, uz, wp, itStart, itEnd, nSteps,
               dt, nSnaps, itSnaps);
  }
  return 0;
}

/*
 * Near field on a Cartesian screen (no counterpart in the reference, whose
 * screen is polar): the observer of index (iX, iY) is (screenX[iX],
 * screenY[iY], L), laid out like the polar screen with x in place of the
 * radius and y in place of phi; everything else is near_point above.
 */
int oracle_near_xy(int comp, double *const *spectra,
                   const double *x, const double *y, const double *z,
                   const double *ux, const double *uy, const double *uz,
                   double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                   const double *omega, const double *screenX, const double *screenY, double distanceToScreen,
                   uint32_t nOmega, uint32_t nX, uint32_t nY,
                   double dt, uint32_t nSnaps, const uint32_t *itSnaps, int nthreads) {
  if (!(comp == OR_TOTAL || comp == OR_CARTESIAN || comp == OR_CARTESIAN_COMPLEX)) return -1;
  if (itEnd == 0 || nSteps == 0) return -2;
  const int64_t nTotal = (int64_t)nX * nY * nOmega;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif

#pragma omp parallel for schedule(static)
  for (int64_t gti = 0; gti < nTotal; gti++) {
    const uint32_t iY = (uint32_t)(gti / ((int64_t)nOmega * nX));
    const uint32_t iX = (uint32_t)((gti - (int64_t)iY * nOmega * nX) / nOmega);
    const uint32_t iOmega = (uint32_t)(gti - (int64_t)iY * nOmega * nX - (int64_t)iX * nOmega);
    const v3 scr = mk(screenX[iX], screenY[iY], distanceToScreen);
    near_point(comp, spectra, gti, nTotal, scr, omega[iOmega], x, y, z, ux, uy, uz, wp, itStart, itEnd, nSteps,
               dt, nSnaps, itSnaps);
  }
  return 0;
}

/*
 * Far-field adjoint, one particle, FP64.  Given G = dL/dS on the grid (one
 * array for `total`, three for `cartesian`, laid out like the spectra), adds
 * dL/d(x, y, z, ux, uy, uz) for every sample of the particle and returns
 * dL/dw in *gw.  S is the one snapshot record (nSnaps must be 1): S = w dt^2
 * |F|^2 with F the sum over the active steps up to the record
 * (kernel_farfield.cl:59-106); the |dphase| < pi mask is held fixed.
 *
 * Chain rule written per (grid point, step):
 *   dL/dA_c   = H_c Re(conj(F_c) e^{i phi}),            H_c = 2 w dt^2 G_c
 *   dL/dphi   = -sum_c H_c A_c Im(conj(F_c) e^{i phi}),  phi = omega (t - n.x)
 *   dL/dx     = -omega n dL/dphi
 *   A = c1 (n - m) - c2 a,  c2 = 1/(1 - n.m),  c1 = (n.a) c2^2
 *   dL/dm     = 2 c1 c2 n ((n - m).gA) - c1 gA - c2^2 n (a.gA)
 *   dL/da     = c2^2 n ((n - m).gA) - c2 gA
 *   m = (b_s + b_{s+1})/2,  a = (b_{s+1} - b_s)/dt,  b = u/sqrt(1 + u.u)
 *   dL/du     = (gb - b (b.gb)) sqrt(1 - b.b)
 * gx..guz (nSteps each) are accumulated into, not overwritten.
 */
int oracle_far_adjoint(int comp, const double *const *G,
                       const double *x, const double *y, const double *z,
                       const double *ux, const double *uy, const double *uz,
                       double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                       const double *omega, const double *sinTheta, const double *cosTheta,
                       const double *sinPhi, const double *cosPhi,
                       uint32_t nOmega, uint32_t nTheta, uint32_t nPhi,
                       double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                       double *gx, double *gy, double *gz,
                       double *gux, double *guy, double *guz, double *gw, int nthreads) {
  if (!(comp == OR_TOTAL || comp == OR_CARTESIAN)) return -1;
  if (itEnd == 0 || nSteps == 0) return -2;
  if (nSnaps != 1) return -4;
  const int64_t nTotal = (int64_t)nTheta * nPhi * nOmega;
  const uint32_t n = nSteps;
  /* record step of the single snapshot (kernel_farfield.cl:53-57, 100) */
  int64_t rec = -1;
  {
    const uint32_t iSnap = first_snap(itStart, nSnaps, itSnaps);
    if (iSnap < nSnaps) {
      const int64_t r = (int64_t)itSnaps[iSnap] - 2 - (int64_t)itStart;
      if (r >= 0 && r <= (int64_t)itEnd - 2) rec = r;
    }
  }
  if (rec < 0) return 0;
  int64_t nAct = (int64_t)n - 1 < (int64_t)itEnd - 1 ? (int64_t)n - 1 : (int64_t)itEnd - 1;
  if (nAct > rec + 1) nAct = rec + 1;
  if (nAct <= 0) return 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = omp_get_max_threads();
#else
  (void)nthreads;
  const int nth = 1;
#endif
  /* per-thread accumulators: dL/dx (3n), dL/db (3n), dL/dw */
  double *acc = (double *)calloc((size_t)nth * (6 * (size_t)n + 1), sizeof(double));
  if (!acc) return -5;
  const double dtInv = 1.0 / dt;

#pragma omp parallel
  {
#ifdef _OPENMP
    const int tno = omp_get_thread_num();
#else
    const int tno = 0;
#endif
    double *ax_ = acc + (size_t)tno * (6 * (size_t)n + 1);
    double *ab_ = ax_ + 3 * (size_t)n;
    double *aw_ = ab_ + 3 * (size_t)n;
#pragma omp for schedule(static)
    for (int64_t gti = 0; gti < nTotal; gti++) {
      const uint32_t iPhi = (uint32_t)(gti / ((int64_t)nOmega * nTheta));
      const uint32_t iTheta = (uint32_t)((gti - (int64_t)iPhi * nOmega * nTheta) / nOmega);
      const uint32_t iOmega = (uint32_t)(gti - (int64_t)iPhi * nOmega * nTheta - (int64_t)iTheta * nOmega);
      const double om = omega[iOmega];
      const v3 nv = mk(sinTheta[iTheta] * cosPhi[iPhi], sinTheta[iTheta] * sinPhi[iPhi], cosTheta[iTheta]);
      double Hc[3];
      if (comp == OR_TOTAL) {
        Hc[0] = Hc[1] = Hc[2] = 2.0 * wp * dt * dt * G[0][gti];
      } else {
        for (int c = 0; c < 3; c++) Hc[c] = 2.0 * wp * dt * dt * G[c][gti];
      }
      /* forward: F over the active steps up to the record */
      double Fr[3] = {0, 0, 0}, Fi[3] = {0, 0, 0};
      double prev = 0.0;
      for (int64_t s = 0; s < nAct; s++) {
        const double t = (double)(uint32_t)(itStart + (uint32_t)s) * dt;
        const double ph = om * (t - dot3(mk(x[s], y[s], z[s]), nv));
        const double d = fabs(ph - prev);
        prev = ph;
        if (!(d < M_PI)) continue;
        v3 b0 = mk(ux[s], uy[s], uz[s]), b1 = mk(ux[s + 1], uy[s + 1], uz[s + 1]);
        b0 = scl3(1.0 / sqrt(1.0 + dot3(b0, b0)), b0);
        b1 = scl3(1.0 / sqrt(1.0 + dot3(b1, b1)), b1);
        const v3 a = scl3(dtInv, sub3(b1, b0));
        const v3 m = scl3(0.5, add3(b1, b0));
        double c2 = 1.0 / (1.0 - dot3(m, nv));
        double c1 = dot3(a, nv) * c2 * c2;
        const v3 A = sub3(scl3(c1, sub3(nv, m)), scl3(c2, a));
        const double cs = cos(ph), sn = sin(ph);
        const double Av[3] = {A.x, A.y, A.z};
        for (int c = 0; c < 3; c++) { Fr[c] += Av[c] * cs; Fi[c] += Av[c] * sn; }
      }
      if (comp == OR_TOTAL)
        aw_[0] += dt * dt * G[0][gti] * (Fr[0] * Fr[0] + Fr[1] * Fr[1] + Fr[2] * Fr[2] +
                                         Fi[0] * Fi[0] + Fi[1] * Fi[1] + Fi[2] * Fi[2]);
      else
        for (int c = 0; c < 3; c++) aw_[0] += dt * dt * G[c][gti] * (Fr[c] * Fr[c] + Fi[c] * Fi[c]);
      /* backward */
      prev = 0.0;
      for (int64_t s = 0; s < nAct; s++) {
        const double t = (double)(uint32_t)(itStart + (uint32_t)s) * dt;
        const double ph = om * (t - dot3(mk(x[s], y[s], z[s]), nv));
        const double d = fabs(ph - prev);
        prev = ph;
        if (!(d < M_PI)) continue;
        v3 b0 = mk(ux[s], uy[s], uz[s]), b1 = mk(ux[s + 1], uy[s + 1], uz[s + 1]);
        b0 = scl3(1.0 / sqrt(1.0 + dot3(b0, b0)), b0);
        b1 = scl3(1.0 / sqrt(1.0 + dot3(b1, b1)), b1);
        const v3 a = scl3(dtInv, sub3(b1, b0));
        const v3 m = scl3(0.5, add3(b1, b0));
        const double c2 = 1.0 / (1.0 - dot3(m, nv));
        const double c1 = dot3(a, nv) * c2 * c2;
        const v3 nm = sub3(nv, m);
        const v3 A = sub3(scl3(c1, nm), scl3(c2, a));
        const double cs = cos(ph), sn = sin(ph);
        const double Av[3] = {A.x, A.y, A.z};
        double gA[3], gph = 0.0;
        for (int c = 0; c < 3; c++) {
          const double qre = Fr[c] * cs + Fi[c] * sn;   /* Re(conj(F) e^{i ph}) */
          const double qim = Fr[c] * sn - Fi[c] * cs;   /* Im(conj(F) e^{i ph}) */
          gA[c] = Hc[c] * qre;
          gph -= Hc[c] * Av[c] * qim;
        }
        const v3 gAv = mk(gA[0], gA[1], gA[2]);
        /* x */
        ax_[3 * s + 0] += -om * nv.x * gph;
        ax_[3 * s + 1] += -om * nv.y * gph;
        ax_[3 * s + 2] += -om * nv.z * gph;
        /* m, a */
        const double nmg = dot3(nm, gAv), ag = dot3(a, gAv);
        const v3 gm = sub3(sub3(scl3(2.0 * c1 * c2 * nmg, nv), scl3(c1, gAv)), scl3(c2 * c2 * ag, nv));
        const v3 ga = sub3(scl3(c2 * c2 * nmg, nv), scl3(c2, gAv));
        /* b_s, b_{s+1} */
        ab_[3 * s + 0] += 0.5 * gm.x - dtInv * ga.x;
        ab_[3 * s + 1] += 0.5 * gm.y - dtInv * ga.y;
        ab_[3 * s + 2] += 0.5 * gm.z - dtInv * ga.z;
        ab_[3 * (s + 1) + 0] += 0.5 * gm.x + dtInv * ga.x;
        ab_[3 * (s + 1) + 1] += 0.5 * gm.y + dtInv * ga.y;
        ab_[3 * (s + 1) + 2] += 0.5 * gm.z + dtInv * ga.z;
      }
    }
  }
  /* reduce threads in order, then b -> u */
  for (int tno = 0; tno < nth; tno++) {
    const double *ax_ = acc + (size_t)tno * (6 * (size_t)n + 1);
    const double *ab_ = ax_ + 3 * (size_t)n;
    for (uint32_t s = 0; s < n; s++) {
      gx[s] += ax_[3 * s];
      gy[s] += ax_[3 * s + 1];
      gz[s] += ax_[3 * s + 2];
    }
    *gw += ab_[3 * (size_t)n];
  }
  for (uint32_t s = 0; s < n; s++) {
    double gb[3] = {0, 0, 0};
    for (int tno = 0; tno < nth; tno++) {
      const double *ab_ = acc + (size_t)tno * (6 * (size_t)n + 1) + 3 * (size_t)n;
      for (int c = 0; c < 3; c++) gb[c] += ab_[3 * s + c];
    }
    const v3 u = mk(ux[s], uy[s], uz[s]);
    const double gInv = 1.0 / sqrt(1.0 + dot3(u, u));
    const v3 b = scl3(gInv, u);
    const double bg = b.x * gb[0] + b.y * gb[1] + b.z * gb[2];
    gux[s] += gInv * (gb[0] - b.x * bg);
    guy[s] += gInv * (gb[1] - b.y * bg);
    guz[s] += gInv * (gb[2] - b.z * bg);
  }
  free(acc);
  return 0;
}
