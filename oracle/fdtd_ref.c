/* FP64 C restatement of the oracle's cavity stepping -- TEST INFRASTRUCTURE
 * ONLY.  Nothing in the product links or calls this file; tests and the
 * bench's CPU-baseline leg use it as the checker at full configuration sizes,
 * where the NumPy restatement would take hours.
 *
 * ceep_oracle_cavity2d follows oracle/fdtd2d.py::run_cavity and
 * ceep_oracle_cavity3d follows oracle/fdtd3d.py::run_cavity statement by
 * statement, in the same floating-point operation order, so results are
 * bitwise those of the NumPy oracle (tests/test_oracle_c.py checks it).  That
 * oracle in turn restates the reference:
 *   2D  src/ceep/solvers/fdtd_2d_batched.py:540-611 (run_cpu form on the box
 *       the inert CPML encloses; H :472-483, E :492-498, inject :509-515,
 *       record :517-526)
 *   3D  src/ceep/solvers/fdtd_3d_batched.py:753-884 (H and E updates),
 *       :949-958 (inject, record)
 *   ADE src/ceep/solvers/fdtd_2d.py:163-178 with dispersive.py:359-404
 *       (E_old = E; E = ca E + cb curl; E -= cb sum(gamma J); J update)
 * Coefficient maps are formed by the NumPy oracle and passed in.
 *
 * Build: oracle/build_oracle.py (gcc -O2 -fopenmp -ffp-contract=off, so no
 * fused multiply-add changes a rounding).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { POLE_DEBYE = 0, POLE_LORENTZ = 1 };

/* One cell's ADE-composed E update (fdtd_2d.py:163-178, dispersive.py:359-404).
 * j/jp point at pole 0 of this cell; consecutive poles are `pstride` apart.  */
static inline double ade_cell(double e, double ca, double cb, double curl, double dt,
                              int npoles, const int* kind, const double* k1,
                              const double* k2, const double* k3, const double* gam,
                              long long mstride, double* j, double* jp, long long pstride) {
  if (npoles == 0) return ca * e + cb * curl;
  const double e_old = e;
  double sg = gam[0] * j[0];
  for (int q = 1; q < npoles; ++q) sg = sg + gam[q * mstride] * j[q * pstride];
  double en = ca * e + cb * curl;
  en = en - cb * sg;
  const double dd = (en - e_old) / dt;
  for (int q = 0; q < npoles; ++q) {
    const long long m = q * mstride, c = q * pstride;
    if (kind[q] == POLE_DEBYE) {
      j[c] = k1[m] * j[c] + k2[m] * dd;
    } else {
      const double jn = k1[m] * j[c] + k3[m] * jp[c] + k2[m] * dd;
      jp[c] = j[c];
      j[c] = jn;
    }
  }
  return en;
}

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

/* 2D TMz cavity on the Mx x My box (oracle/fdtd2d.py::run_cavity).
 * ca, cb: [Mx*My]; db: scalar dt/mu0 (da = 1); pole maps [npoles][Mx*My];
 * src: [B][2] box coordinates (strictly inside); wave: [T] or [B][T];
 * prb: [P][2] box coordinates, (-1, -1) for probes outside the box (they read 0);
 * traces: [B][P][T], overwritten.  Returns 0, or -1 when memory runs out. */
int ceep_oracle_cavity2d(int B, int Mx, int My, int T, double inv, double dt, double db,
                         const double* ca, const double* cb, int npoles, const int* kind,
                         const double* k1, const double* k2, const double* k3,
                         const double* gam, const int* src, const double* wave,
                         int wave_per_b, int P, const int* prb, double* traces, int threads) {
  set_threads(threads);
  const long long plane = (long long)Mx * My;
  const long long nf = (long long)B * plane;
  double* ez = calloc((size_t)nf, sizeof(double));
  double* hx = calloc((size_t)nf, sizeof(double));
  double* hy = calloc((size_t)nf, sizeof(double));
  double* J = npoles ? calloc((size_t)(nf * npoles), sizeof(double)) : NULL;
  double* Jp = npoles ? calloc((size_t)(nf * npoles), sizeof(double)) : NULL;
  if (!ez || !hx || !hy || (npoles && (!J || !Jp))) {
    free(ez); free(hx); free(hy); free(J); free(Jp);
    return -1;
  }
  memset(traces, 0, sizeof(double) * (size_t)B * P * T);
  const double dbi = db * inv;
  for (int step = 0; step < T; ++step) {
    /* H: hx[:, :, :-1] and hy[:, :-1, :] (fdtd2d.py run_cavity) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int i = 0; i < Mx; ++i) {
        const long long r = b * plane + (long long)i * My;
        for (int j = 0; j < My - 1; ++j) hx[r + j] = hx[r + j] - dbi * (ez[r + j + 1] - ez[r + j]);
        if (i < Mx - 1)
          for (int j = 0; j < My; ++j) hy[r + j] = hy[r + j] + dbi * (ez[r + My + j] - ez[r + j]);
      }
    /* E (+ ADE) on the interior [1, Mx-1) x [1, My-1) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int i = 1; i < Mx - 1; ++i) {
        const long long r = b * plane + (long long)i * My;
        const long long mr = (long long)i * My;
        for (int j = 1; j < My - 1; ++j) {
          const double curl = (hy[r + j] - hy[r - My + j]) * inv - (hx[r + j] - hx[r + j - 1]) * inv;
          ez[r + j] = ade_cell(ez[r + j], ca[mr + j], cb[mr + j], curl, dt, npoles, kind,
                               k1 + mr + j, k2 + mr + j, k3 + mr + j, gam + mr + j, plane,
                               J ? J + r + j : NULL, Jp ? Jp + r + j : NULL, nf);
        }
      }
    for (int b = 0; b < B; ++b) {
      const double w = wave_per_b ? wave[(long long)b * T + step] : wave[step];
      ez[b * plane + (long long)src[2 * b] * My + src[2 * b + 1]] += w;
    }
    for (int p = 0; p < P; ++p) {
      if (prb[2 * p] < 0) continue;
      const long long c = (long long)prb[2 * p] * My + prb[2 * p + 1];
      for (int b = 0; b < B; ++b) traces[((long long)b * P + p) * T + step] = ez[b * plane + c];
    }
  }
  free(ez); free(hx); free(hy); free(J); free(Jp);
  return 0;
}

/* 3D cavity on the mx x my x mz box (oracle/fdtd3d.py::run_cavity).
 * ca, cb, da, db: [mx*my*mz]; pole maps [npoles][cells]; src: [B][3];
 * prb: [P][3] ((-1, -1, -1) outside); traces: [B][P][T] of Ez. */
int ceep_oracle_cavity3d(int B, int mx, int my, int mz, int T, double inv, double dt,
                         const double* ca, const double* cb, const double* da,
                         const double* db, int npoles, const int* kind, const double* k1,
                         const double* k2, const double* k3, const double* gam,
                         const int* src, const double* wave, int wave_per_b, int P,
                         const int* prb, double* traces, int threads) {
  set_threads(threads);
  const long long sy = mz, sx = (long long)my * mz;
  const long long vol = (long long)mx * sx;
  const long long nf = (long long)B * vol;
  double* f[6];
  int ok = 1;
  for (int c = 0; c < 6; ++c) ok &= (f[c] = calloc((size_t)nf, sizeof(double))) != NULL;
  double* J[3] = {NULL, NULL, NULL};
  double* Jp[3] = {NULL, NULL, NULL};
  for (int c = 0; c < 3 && npoles; ++c) {
    ok &= (J[c] = calloc((size_t)(nf * npoles), sizeof(double))) != NULL;
    ok &= (Jp[c] = calloc((size_t)(nf * npoles), sizeof(double))) != NULL;
  }
  if (!ok) {
    for (int c = 0; c < 6; ++c) free(f[c]);
    for (int c = 0; c < 3; ++c) { free(J[c]); free(Jp[c]); }
    return -1;
  }
  double *ex = f[0], *ey = f[1], *ez = f[2], *hx = f[3], *hy = f[4], *hz = f[5];
  memset(traces, 0, sizeof(double) * (size_t)B * P * T);
  for (int step = 0; step < T; ++step) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int i = 0; i < mx; ++i)
        for (int j = 0; j < my; ++j) {
          const long long m0 = i * sx + j * sy;
          const long long r = b * vol + m0;
          if (j < my - 1)   /* hx[:, :, :my-1, :mz-1] */
            for (int k = 0; k < mz - 1; ++k)
              hx[r + k] = da[m0 + k] * hx[r + k] +
                          db[m0 + k] * ((ey[r + k + 1] - ey[r + k]) * inv -
                                        (ez[r + sy + k] - ez[r + k]) * inv);
          if (i < mx - 1)   /* hy[:, :mx-1, :, :mz-1] */
            for (int k = 0; k < mz - 1; ++k)
              hy[r + k] = da[m0 + k] * hy[r + k] +
                          db[m0 + k] * ((ez[r + sx + k] - ez[r + k]) * inv -
                                        (ex[r + k + 1] - ex[r + k]) * inv);
          if (i < mx - 1 && j < my - 1)   /* hz[:, :mx-1, :my-1, :] */
            for (int k = 0; k < mz; ++k)
              hz[r + k] = da[m0 + k] * hz[r + k] +
                          db[m0 + k] * ((ex[r + sy + k] - ex[r + k]) * inv -
                                        (ey[r + sx + k] - ey[r + k]) * inv);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
      for (int i = 1; i < mx - 1; ++i)
        for (int j = 1; j < my - 1; ++j) {
          const long long m0 = i * sx + j * sy;
          const long long r = b * vol + m0;
          for (int k = 1; k < mz - 1; ++k) {
            const long long m = m0 + k, c = r + k;
            const double cx = (hz[c] - hz[c - sy]) * inv - (hy[c] - hy[c - 1]) * inv;
            const double cy = (hx[c] - hx[c - 1]) * inv - (hz[c] - hz[c - sx]) * inv;
            const double cz = (hy[c] - hy[c - sx]) * inv - (hx[c] - hx[c - sy]) * inv;
            const double* ce[3] = {&cx, &cy, &cz};
            for (int comp = 0; comp < 3; ++comp) {
              double* E = f[comp];
              E[c] = ade_cell(E[c], ca[m], cb[m], *ce[comp], dt, npoles, kind, k1 + m, k2 + m,
                              k3 + m, gam + m, vol, J[comp] ? J[comp] + c : NULL,
                              Jp[comp] ? Jp[comp] + c : NULL, nf);
            }
          }
        }
    for (int b = 0; b < B; ++b) {
      const double w = wave_per_b ? wave[(long long)b * T + step] : wave[step];
      ez[b * vol + src[3 * b] * sx + src[3 * b + 1] * sy + src[3 * b + 2]] += w;
    }
    for (int p = 0; p < P; ++p) {
      if (prb[3 * p] < 0) continue;
      const long long c = prb[3 * p] * sx + prb[3 * p + 1] * sy + prb[3 * p + 2];
      for (int b = 0; b < B; ++b) traces[((long long)b * P + p) * T + step] = ez[b * vol + c];
    }
  }
  for (int c = 0; c < 6; ++c) free(f[c]);
  for (int c = 0; c < 3; ++c) { free(J[c]); free(Jp[c]); }
  return 0;
}
