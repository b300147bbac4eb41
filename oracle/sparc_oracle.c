/*
 * sparc_oracle.c -- FP64 CPU restatement of the SPARC forward/adjoint
 * time-stepping path.  TEST INFRASTRUCTURE ONLY: this file is the checker the
 * tests, the smoke check and the CPU baseline of the benchmark compare the
 * GPU engine against.  It is never linked into or called by the product.
 *
 * Every function cites the reference file:line it restates.  Arrays are
 * Fortran column-major (nx,ny,nz), index i + nx*(j + ny*k) with 0-based
 * i,j,k; plane numbers handed in from outside (e_rad, o_rad) are 1-based.
 *
 * Single-precision literals of the reference (no -r8, makefile:27) are kept
 * through SP(x) = (double)(float)(x), e.g. pi (initialize.f90:38).
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   dbyd2/mtridag   bit-level vs the reference's own dbyd2.f90+mtridag.f90
 *                   built into oracle/_ref, and vs a dense numpy solve;
 *   fft/derivs      numpy.fft and exact spectral derivatives of trig
 *                   polynomials;
 *   LDDRK step      closed-form amplification polynomial on y'=lambda*y;
 *   z filter        constants -> identity, (-1)^k -> 0, brute force stencil;
 *   Lagrange interp exact on degree-6 polynomials;
 *   compute_tt      exact shift recovery on a quadratic cross-correlation;
 *   composed RHS    plane-wave dispersion in a uniform periodic medium,
 *                   2-D (option 17) vs 3-D (option 4) relations, and the
 *                   GPU engine at FP64.
 */
#include <complex.h>
#include <omp.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sparc_gpu.h"

/* complex.h defines I; the restated Fortran loops use I as an index */
#undef I
#define IMAG _Complex_I

#define SP(x) ((double)(float)(x))
typedef double complex cplx;

/* initialize.f90:37-38: rsun, pi are REAL*8 parameters set from default-real
 * literals, i.e. rounded to single precision. */
#define PI_R SP(3.14159265358979)
#define RSUN_R SP(69598946770.0)

static inline long IDX(int i, int j, int k, int nx, int ny) {
  return (long)i + (long)nx * ((long)j + (long)ny * (long)k);
}

/* ======================================================================
 * Compact sixth-order z derivative: DBYD2 (dbyd2.f90:8-237) with MTRIDAG
 * (mtridag.f90:7-114).  A(N1,N2) = d/d(index/(N2-1)) of B along the 2nd
 * index.  ibc 0..4: ispecial=0 tridiagonal; ibc 5: periodic (ispecial=1).
 * ====================================================================== */

/* MTRIDAG ispecial=0 (mtridag.f90:94-111) on column stride n1. */
static void mtridag0(const double *a, const double *b, const double *c,
                     double *u, long n1, int n2) {
  double *betad = malloc(sizeof(double) * n2), *gam = malloc(sizeof(double) * n2);
  betad[0] = b[0];
  for (int j = 1; j < n2; j++) {
    gam[j] = c[j - 1] / betad[j - 1];
    betad[j] = b[j] - a[j] * gam[j];
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n1; i++) {
    double ud_prev = u[i] / betad[0];
    u[i] = ud_prev;
    for (int j = 1; j < n2; j++) {
      double ud = (u[i + n1 * j] - a[j] * ud_prev) / betad[j];
      u[i + n1 * j] = ud;
      ud_prev = ud;
    }
    for (int j = n2 - 2; j >= 0; j--)
      u[i + n1 * j] = u[i + n1 * j] - gam[j + 1] * u[i + n1 * (j + 1)];
  }
  free(betad);
  free(gam);
}

/* MTRIDAG ispecial=1, periodic (mtridag.f90:19-65). */
static void mtridag1(const double *a, const double *b, const double *c,
                     double *u, long n1, int n2) {
  double *betad = malloc(sizeof(double) * n2), *betas = malloc(sizeof(double) * n2);
  double *betau = malloc(sizeof(double) * n2), *alpha = malloc(sizeof(double) * n2);
  double *alphab = malloc(sizeof(double) * n2);
  for (int i = 0; i < n2; i++) {
    betad[i] = b[i];
    betas[i] = 0;
    betau[i] = c[i];
    alpha[i] = a[i];
    alphab[i] = 0;
  }
  betas[0] = a[0];
  betas[n2 - 2] = c[n2 - 2];
  alphab[0] = c[n2 - 1];
  alphab[n2 - 2] = a[n2 - 1];
  for (int i = 1; i < n2 - 1; i++) {
    betad[i] = betad[i] - alpha[i - 1] * betau[i - 1];
    betas[i] = betas[i] - alpha[i - 1] * betas[i - 1];
    alpha[i] = alpha[i] / betad[i];
    alphab[i] = (alphab[i] - alphab[i - 1] * betau[i - 1]) / betad[i];
  }
  for (int i = 0; i < n2 - 1; i++) betad[n2 - 1] = betad[n2 - 1] - alphab[i] * betas[i];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n1; i++) {
    double *ud = malloc(sizeof(double) * n2);
    ud[0] = u[i];
    for (int j = 1; j < n2 - 1; j++) ud[j] = u[i + n1 * j] - alpha[j - 1] * ud[j - 1];
    ud[n2 - 1] = u[i + n1 * (n2 - 1)];
    for (int j = 0; j < n2 - 1; j++) ud[n2 - 1] = ud[n2 - 1] - alphab[j] * ud[j];
    ud[n2 - 1] = ud[n2 - 1] / betad[n2 - 1];
    ud[n2 - 2] = (ud[n2 - 2] - betau[n2 - 2] * ud[n2 - 1]) / betad[n2 - 2];
    for (int j = n2 - 3; j >= 0; j--)
      ud[j] = (ud[j] - betau[j] * ud[j + 1] - betas[j] * ud[n2 - 1]) / betad[j];
    for (int j = 0; j < n2; j++) u[i + n1 * j] = ud[j];
    free(ud);
  }
  free(betad);
  free(betas);
  free(betau);
  free(alpha);
  free(alphab);
}

/* DBYD2 (dbyd2.f90:8-237).  On entry A holds the caller-set boundary values
 * where the reference expects them (ibc 2: A(:,1); ibc 3: A(:,N2); ibc 4:
 * both). */
int orc_dbyd2(double *A, const double *B, long n1, int n2, int ibc) {
  if (ibc < 0 || ibc > 5 || n2 < 5) return SPARC_EINVAL;
  double d2i = (ibc == 5) ? (double)n2 : (double)(n2 - 1);
  double *uppr = calloc(n2, sizeof(double)), *diag = malloc(sizeof(double) * n2);
  double *lowr = calloc(n2, sizeof(double));
  for (int j = 0; j < n2; j++) diag[j] = 1.0;
#define A_(i, j) A[(i) + n1 * (long)(j)]
#define B_(i, j) B[(i) + n1 * (long)(j)]
  double C1 = d2i * 7.0 / 9.0, C2 = d2i / 36.0;
  if (ibc < 5) {
#pragma omp parallel for schedule(static)
    for (int I = 2; I < n2 - 2; I++)
      for (long J = 0; J < n1; J++)
        A_(J, I) = C1 * (B_(J, I + 1) - B_(J, I - 1)) + C2 * (B_(J, I + 2) - B_(J, I - 2));
    /* fifth order at rows 2 and N2-1 (dbyd2.f90:103-118) */
    C1 = SP(-43. / 96.) * d2i;
    C2 = SP(-5. / 6.) * d2i;
    double C3 = 9. / 8. * d2i;
    double C4 = SP(1. / 6.) * d2i;
    double C5 = SP(-1. / 96.) * d2i;
    for (long I = 0; I < n1; I++)
      A_(I, 1) = C1 * B_(I, 0) + C2 * B_(I, 1) + C3 * B_(I, 2) + C4 * B_(I, 3) + C5 * B_(I, 4);
    for (long I = 0; I < n1; I++)
      A_(I, n2 - 2) = -(C1 * B_(I, n2 - 1) + C2 * B_(I, n2 - 2) + C3 * B_(I, n2 - 3) +
                        C4 * B_(I, n2 - 4) + C5 * B_(I, n2 - 5));
    /* fourth order at rows 1 and N2 (dbyd2.f90:121-146) */
    C1 = SP(-10 / 3.) * d2i;
    C2 = -3.0 * d2i;
    C3 = 6. * d2i;
    C4 = SP(1. / 3.) * d2i;
    if (ibc == 0) {
      for (long I = 0; I < n1; I++) A_(I, 0) = 0.0;
    } else if (ibc == 1 || ibc == 3) {
      for (long I = 0; I < n1; I++)
        A_(I, 0) = C1 * B_(I, 0) + C2 * B_(I, 1) + C3 * B_(I, 2) + C4 * B_(I, 3);
    }
    if (ibc == 0) {
      for (long I = 0; I < n1; I++) A_(I, n2 - 1) = 0.0;
    } else if (ibc == 1 || ibc == 2) {
      for (long I = 0; I < n1; I++)
        A_(I, n2 - 1) = -(C1 * B_(I, n2 - 1) + C2 * B_(I, n2 - 2) + C3 * B_(I, n2 - 3) +
                          C4 * B_(I, n2 - 4));
    }
    /* matrix and preconditioning (dbyd2.f90:151-196) */
    double alpha = SP(1. / 3.);
    for (int J = 2; J < n2 - 2; J++) uppr[J] = alpha;
    double alpha2 = 3. / 4., gamma2 = 1. / 8., alpha1 = 6.0, beta1 = 3.;
    double cst = 1. / (alpha2 - beta1 * gamma2);
    double up1 = (alpha1 * alpha2 - beta1) * cst;
    if (ibc % 2 != 0)
      for (long i = 0; i < n1; i++) A_(i, 0) = (alpha2 * A_(i, 0) - beta1 * A_(i, 1)) * cst;
    if (ibc == 1 || ibc == 2)
      for (long i = 0; i < n1; i++)
        A_(i, n2 - 1) = (alpha2 * A_(i, n2 - 1) - beta1 * A_(i, n2 - 2)) * cst;
    uppr[0] = (ibc % 2 == 0) ? 0.0 : up1;
    uppr[1] = alpha2;
    uppr[n2 - 2] = gamma2;
    /* uppr(N2) is never set by the reference and LOWR(1)=UPPR(N2) is never
     * read by MTRIDAG; it is left 0 here. */
    for (int I = 0; I < n2; I++) lowr[I] = uppr[n2 - 1 - I];
    lowr[n2 - 1] = (ibc != 1 && ibc != 2) ? 0.0 : up1;
    mtridag0(lowr, diag, uppr, A, n1, n2);
  } else {
    for (long J = 0; J < n1; J++) {
      A_(J, 1) = C1 * (B_(J, 2) - B_(J, 0)) + C2 * (B_(J, 3) - B_(J, n2 - 1));
      A_(J, 0) = C1 * (B_(J, 1) - B_(J, n2 - 1)) + C2 * (B_(J, 2) - B_(J, n2 - 2));
      A_(J, n2 - 2) = C1 * (B_(J, n2 - 1) - B_(J, n2 - 3)) + C2 * (B_(J, 0) - B_(J, n2 - 4));
      A_(J, n2 - 1) = C1 * (B_(J, 0) - B_(J, n2 - 2)) + C2 * (B_(J, 1) - B_(J, n2 - 3));
    }
#pragma omp parallel for schedule(static)
    for (int I = 2; I < n2 - 2; I++)
      for (long J = 0; J < n1; J++)
        A_(J, I) = C1 * (B_(J, I + 1) - B_(J, I - 1)) + C2 * (B_(J, I + 2) - B_(J, I - 2));
    double alpha = SP(1. / 3.);
    for (int J = 0; J < n2; J++) uppr[J] = lowr[J] = alpha;
    mtridag1(lowr, diag, uppr, A, n1, n2);
  }
#undef A_
#undef B_
  free(uppr);
  free(diag);
  free(lowr);
  return SPARC_OK;
}

/* ======================================================================
 * FFT: recursive mixed-radix (4,2,3,5) complex DFT in FP64.  Stands in for
 * FFTW 3.3.4's r2c/c2r (makefile:29), whose rounding is not reproducible;
 * the tests pin it against numpy.fft.
 * ====================================================================== */
typedef struct {
  int n;
  cplx *w; /* w[m] = exp(-2 pi i m / n) */
} fftplan;

static fftplan *plan_new(int n) {
  fftplan *p = malloc(sizeof(fftplan));
  p->n = n;
  p->w = malloc(sizeof(cplx) * (n > 0 ? n : 1));
  for (int m = 0; m < n; m++) {
    double ang = -2.0 * M_PI * (double)m / (double)n;
    p->w[m] = cos(ang) + IMAG * sin(ang);
  }
  return p;
}
static void plan_free(fftplan *p) {
  if (!p) return;
  free(p->w);
  free(p);
}

static void fft_rec(const cplx *in, cplx *out, int n, int stride, const fftplan *pl,
                    int sign) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
  int m = n / p;
  for (int r = 0; r < p; r++) fft_rec(in + (long)r * stride, out + r * m, m, stride * p, pl, sign);
  int N0 = pl->n, tstep = N0 / n;
  cplx t[8], *tt = (p <= 8) ? t : malloc(sizeof(cplx) * p);
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) {
      cplx wv = pl->w[((long)r * k * tstep) % N0];
      if (sign > 0) wv = conj(wv);
      tt[r] = out[r * m + k] * wv;
    }
    for (int q = 0; q < p; q++) {
      cplx s = 0;
      for (int r = 0; r < p; r++) {
        cplx wv = pl->w[(((long)r * q) % p) * (N0 / p)];
        if (sign > 0) wv = conj(wv);
        s += tt[r] * wv;
      }
      out[q * m + k] = s;
    }
  }
  if (tt != t) free(tt);
}

/* forward (sign -1) or unnormalized inverse (sign +1) complex DFT */
void orc_fft(const double *in_ri, double *out_ri, int n, int sign) {
  fftplan *pl = plan_new(n);
  cplx *in = malloc(sizeof(cplx) * n), *out = malloc(sizeof(cplx) * n);
  for (int m = 0; m < n; m++) in[m] = in_ri[2 * m] + IMAG * in_ri[2 * m + 1];
  fft_rec(in, out, n, 1, pl, sign);
  for (int m = 0; m < n; m++) {
    out_ri[2 * m] = creal(out[m]);
    out_ri[2 * m + 1] = cimag(out[m]);
  }
  free(in);
  free(out);
  plan_free(pl);
}

/* Applies a full-spectrum multiplier along axis 0 (x) or 1 (y) of a 3-D
 * array: out = IDFT_normalized(mult * DFT(in)), real part. */
static void spectral_axis(const double *in, double *out, int nx, int ny, int nz, int axis,
                          const cplx *mult) {
  int n = axis == 0 ? nx : ny;
  long nl = axis == 0 ? (long)ny * nz : (long)nx * nz;
  fftplan *pl = plan_new(n);
#pragma omp parallel
  {
    cplx *b = malloc(sizeof(cplx) * n), *s = malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
    for (long l = 0; l < nl; l++) {
      long off, st;
      if (axis == 0) {
        off = l * nx;
        st = 1;
      } else {
        long i = l % nx, k = l / nx;
        off = i + (long)nx * ny * k;
        st = nx;
      }
      for (int m = 0; m < n; m++) b[m] = in[off + st * m];
      fft_rec(b, s, n, 1, pl, -1);
      for (int m = 0; m < n; m++) s[m] *= mult[m];
      fft_rec(s, b, n, 1, pl, +1);
      for (int m = 0; m < n; m++) out[off + st * m] = creal(b[m]) / (double)n;
    }
    free(b);
    free(s);
  }
  plan_free(pl);
}

/* ddx/ddy with PERIODIC and USE_FFT (derivatives.f90:28-58, 62-170):
 * c2r(eyek * r2c(f)) * stretch, eyek(m) = m*i*norm*2*pi (initialize.f90:204-210),
 * the Nyquist bin being discarded by c2r. */
static void deriv_mult(cplx *mult, int n, double stretch) {
  for (int m = 0; m < n; m++) {
    int ms = (m <= n / 2) ? m : m - n;
    if (2 * m == n) ms = 0;
    mult[m] = IMAG * (double)ms * 2.0 * PI_R * stretch;
  }
}

void orc_ddx(const double *in, double *out, int nx, int ny, int nz, double stretchx) {
  cplx *mult = malloc(sizeof(cplx) * nx);
  deriv_mult(mult, nx, stretchx);
  spectral_axis(in, out, nx, ny, nz, 0, mult);
  free(mult);
}

void orc_ddy(const double *in, double *out, int nx, int ny, int nz, double stretchy) {
  cplx *mult = malloc(sizeof(cplx) * ny);
  deriv_mult(mult, ny, stretchy);
  spectral_axis(in, out, nx, ny, nz, 1, mult);
  free(mult);
}

/* ddz (derivatives.f90:10-24): dbyd2 along z, times stretch(k). */
void orc_ddz(const double *in, double *out, int nx, int ny, int nz, int bc,
             const double *stretch) {
  long n1 = (long)nx * ny;
  orc_dbyd2(out, in, n1, nz, bc);
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < n1; c++) out[c + n1 * k] = out[c + n1 * k] * stretch[k];
}

/* Spectral Fermi low-pass along an axis with decay_coeff (which folds the
 * 1/n of the unnormalized c2r), filter_vars_3d/2d (physics.f90:796-830). */
static void lowpass_axis(double *f, int nx, int ny, int nz, int axis, const double *decay) {
  int n = axis == 0 ? nx : ny;
  cplx *mult = malloc(sizeof(cplx) * n);
  for (int m = 0; m < n; m++) {
    int mm = (m <= n / 2) ? m : n - m;
    mult[m] = decay[mm] * (double)n;
  }
  spectral_axis(f, f, nx, ny, nz, axis, mult);
  free(mult);
}

/* The Fermi low-pass of FILTER_VARS_3D along one axis (physics.f90:796-828),
 * exported for its closed-form check. */
void orc_lowpass(double *f, int nx, int ny, int nz, int axis, const double *decay) {
  lowpass_axis(f, nx, ny, nz, axis, decay);
}

/* ======================================================================
 * The solver context (module globals of initialize.f90:24-157)
 * ====================================================================== */
typedef struct orc {
  sparc_config cfg;
  int nx, ny, nz, nzpml;
  long N, NP;
  double *a, *temp, *scr;             /* 6*N each */
  double *pmlvars, *psipml, *scrpml;  /* npv*NP each */
  int npv;                            /* psi fields: 2, or 6 for the MHD options */
  int nf;                             /* dimfive: 5, 6, or 8 for option 8 */
  double *mw[12];                     /* option 10 work arrays */
  double *bg[SPARC_BG_COUNT];         /* N each (z-profiles expanded) */
  double *prof[SPARC_PROF_COUNT];     /* nz each */
  double *damping_rates;              /* (nx/2+1)*ny */
  double *decay_x, *decay_y;
  double betas[5];
  /* work arrays */
  double *dxdx, *dydy, *dzdz, *div, *p, *rho, *gpx, *gpy, *gpz, *term;
  /* source */
  int src_kind, src_nt, src_active_until;
  int *src_planes;  /* per-pixel 0-based source planes (MHD), NULL = e_rad/o_rad */
  double *src_dw;   /* delta_width_2d */
  double *src_mask, *src_series;
  double *forcing; /* nx*ny */
  int time_old;
  double x0, x1, x2, x3, x4, x5, x6;
  double *LC[7];
  /* kernels */
  int nslots;
  double *snap; /* nslots * 6 * N: xi_x, xi_y, xi_z, acc_x, acc_y, acc_z */
  double *kern[SPARC_KERN_COUNT];
  double *grav;
  int time;
} orc;

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

orc *orc_create(const sparc_config *cfg) {
  if (cfg->nx < 2 || cfg->ny < 1 || cfg->nz < 12) return NULL;
  orc *o = zalloc(sizeof(orc));
  o->cfg = *cfg;
  o->nx = cfg->nx;
  o->ny = cfg->ny;
  o->nz = cfg->nz;
  o->nzpml = cfg->npmlbot + cfg->npmltop;
  o->N = (long)o->nx * o->ny * o->nz;
  o->NP = (long)o->nx * o->ny * (o->nzpml > 0 ? o->nzpml : 1);
  long N = o->N;
  o->nf = (cfg->option == SPARC_OPT_QUIET_PML_3D) ? 5 : (cfg->option == SPARC_OPT_MHD_PML_3D ? 8 : 6);
  o->a = zalloc(sizeof(double) * SPARC_MAXFIELDS * N);
  o->temp = zalloc(sizeof(double) * SPARC_MAXFIELDS * N);
  o->scr = zalloc(sizeof(double) * SPARC_MAXFIELDS * N);
  int mhd = cfg->option == SPARC_OPT_MHD_PML_3D_DISPL || cfg->option == SPARC_OPT_MHD_PML_2D_DISPL ||
            cfg->option == SPARC_OPT_MHD_PML_3D;
  o->npv = mhd ? 6 : 2;
  o->pmlvars = zalloc(sizeof(double) * o->npv * o->NP);
  o->psipml = zalloc(sizeof(double) * o->npv * o->NP);
  o->scrpml = zalloc(sizeof(double) * o->npv * o->NP);
  for (int q = 0; q < 12; q++)
    o->mw[q] = mhd ? zalloc(sizeof(double) * N) : NULL;
  for (int b = 0; b < SPARC_BG_COUNT; b++) o->bg[b] = zalloc(sizeof(double) * N);
  for (int b = 0; b < SPARC_PROF_COUNT; b++) o->prof[b] = zalloc(sizeof(double) * o->nz);
  o->damping_rates = zalloc(sizeof(double) * (o->nx / 2 + 1) * o->ny);
  o->decay_x = zalloc(sizeof(double) * (o->nx / 2 + 1));
  o->decay_y = zalloc(sizeof(double) * (o->ny / 2 + 1));
  double **w[] = {&o->dxdx, &o->dydy, &o->dzdz, &o->div, &o->p,
                  &o->rho,  &o->gpx,  &o->gpy,  &o->gpz, &o->term};
  for (unsigned q = 0; q < sizeof(w) / sizeof(w[0]); q++) *w[q] = zalloc(sizeof(double) * N);
  o->forcing = zalloc(sizeof(double) * o->nx * o->ny);
  for (int q = 0; q < 7; q++) o->LC[q] = zalloc(sizeof(double) * o->nx * o->ny);
  for (int q = 0; q < SPARC_KERN_COUNT; q++) o->kern[q] = zalloc(sizeof(double) * N);
  o->grav = zalloc(sizeof(double) * o->nz);
  /* INITIALIZE_STEP (step.f90:24-41) */
  double optimals[5] = {1.0, 0.5, 0.1665579725791151184, 0.03950410250586729987,
                        0.007810706393337838236};
  /* optimals(1..2) are exact; (3..5) are default-real literals */
  optimals[2] = SP(0.1665579725791151184);
  optimals[3] = SP(0.03950410250586729987);
  optimals[4] = SP(0.007810706393337838236);
  o->betas[4] = optimals[1];
  o->betas[3] = optimals[2] / o->betas[4];
  o->betas[2] = optimals[3] / (o->betas[4] * o->betas[3]);
  o->betas[1] = optimals[4] / (o->betas[4] * o->betas[3] * o->betas[2]);
  o->betas[0] = 0.0;
  for (int s = 0; s < 5; s++) o->betas[s] *= cfg->deltat;
  return o;
}

void orc_destroy(orc *o) {
  if (!o) return;
  free(o->a);
  free(o->temp);
  free(o->scr);
  free(o->src_planes);
  free(o->src_dw);
  free(o->pmlvars);
  free(o->psipml);
  free(o->scrpml);
  for (int q = 0; q < 12; q++) free(o->mw[q]);
  for (int b = 0; b < SPARC_BG_COUNT; b++) free(o->bg[b]);
  for (int b = 0; b < SPARC_PROF_COUNT; b++) free(o->prof[b]);
  free(o->damping_rates);
  free(o->decay_x);
  free(o->decay_y);
  free(o->dxdx);
  free(o->dydy);
  free(o->dzdz);
  free(o->div);
  free(o->p);
  free(o->rho);
  free(o->gpx);
  free(o->gpy);
  free(o->gpz);
  free(o->term);
  free(o->forcing);
  for (int q = 0; q < 7; q++) free(o->LC[q]);
  for (int q = 0; q < SPARC_KERN_COUNT; q++) free(o->kern[q]);
  free(o->grav);
  free(o->snap);
  free(o->src_mask);
  free(o->src_series);
  free(o);
}

double orc_beta(const orc *o, int s) { return o->betas[s]; }

/* threads the OpenMP loops of the oracle run on */
int orc_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

int orc_set_background(orc *o, int id, const double *h, long n) {
  if (id < 0 || id >= SPARC_BG_COUNT) return SPARC_EINVAL;
  long nxy = (long)o->nx * o->ny;
  if (n == o->N) {
    memcpy(o->bg[id], h, sizeof(double) * n);
  } else if (n == o->nz) {
    for (int k = 0; k < o->nz; k++)
      for (long c = 0; c < nxy; c++) o->bg[id][c + nxy * k] = h[k];
  } else {
    return SPARC_EINVAL;
  }
  return SPARC_OK;
}

int orc_set_profile(orc *o, int id, const double *h, int n) {
  if (id < 0 || id >= SPARC_PROF_COUNT || n != o->nz) return SPARC_EINVAL;
  memcpy(o->prof[id], h, sizeof(double) * n);
  return SPARC_OK;
}

int orc_set_damping(orc *o, const double *r, long n) {
  if (n != (long)(o->nx / 2 + 1) * o->ny) return SPARC_EINVAL;
  memcpy(o->damping_rates, r, sizeof(double) * n);
  return SPARC_OK;
}

int orc_set_filter(orc *o, const double *dx, const double *dy) {
  memcpy(o->decay_x, dx, sizeof(double) * (o->nx / 2 + 1));
  memcpy(o->decay_y, dy, sizeof(double) * (o->ny / 2 + 1));
  return SPARC_OK;
}

/* erad_2d / orad_2d and delta_width_2d (physics.f90:66-90) */
int orc_set_source_planes(orc *o, const int32_t *k_xy, long n) {
  free(o->src_planes);
  free(o->src_dw);
  o->src_planes = NULL;
  o->src_dw = NULL;
  if (n == 0) return SPARC_OK;
  long nxy = (long)o->nx * o->ny;
  if (n != nxy) return SPARC_EINVAL;
  o->src_planes = malloc(sizeof(int) * nxy);
  o->src_dw = malloc(sizeof(double) * nxy);
  const double *z = o->prof[SPARC_PROF_Z];
  for (long c = 0; c < nxy; c++) {
    int k = k_xy[c];
    if (k < 2 || k > o->nz - 1) return SPARC_EINVAL;
    o->src_planes[c] = k - 1;
    o->src_dw[c] = 2.0 / (z[k] - z[k - 2]);
  }
  return SPARC_OK;
}

int orc_set_source(orc *o, int kind, const double *mask, long nmask, const double *series,
                   int nt, int active_until) {
  free(o->src_mask);
  free(o->src_series);
  o->src_mask = NULL;
  o->src_series = NULL;
  o->src_kind = kind;
  o->src_nt = nt;
  o->src_active_until = active_until;
  long nxy = (long)o->nx * o->ny;
  if (kind == SPARC_SRC_SERIES) {
    if (nmask != nxy) return SPARC_EINVAL;
    o->src_mask = malloc(sizeof(double) * nxy);
    memcpy(o->src_mask, mask, sizeof(double) * nxy);
    o->src_series = malloc(sizeof(double) * nt);
    memcpy(o->src_series, series, sizeof(double) * nt);
  } else if (kind == SPARC_SRC_LAGRANGE) {
    if (nmask != nxy * nt) return SPARC_EINVAL;
    o->src_mask = malloc(sizeof(double) * nmask);
    memcpy(o->src_mask, mask, sizeof(double) * nmask);
  }
  o->time_old = 0;
  return SPARC_OK;
}

/* LAGRANGE_INTERP (physics.f90:678-737), vr(:,:,t) = mask[(t-1)*nxy ...]. */
static void lagrange_interp(orc *o) {
  long nxy = (long)o->nx * o->ny;
  double timestep = o->cfg.timestep, cad = o->cfg.cadforcing;
  int num_steps = o->cfg.num_steps;
  double xt = (double)o->time * timestep / cad;
  int time1 = (int)floor(xt) + 1;
  xt = o->time * timestep;
  if (o->time_old < time1) {
    o->x0 = (time1 - 4) * cad;
    o->x1 = (time1 - 3) * cad;
    o->x2 = (time1 - 2) * cad;
    o->x3 = (time1 - 1) * cad;
    o->x4 = (time1)*cad;
    o->x5 = (time1 + 1) * cad;
    o->x6 = (time1 + 2) * cad;
    double x0 = o->x0, x1 = o->x1, x2 = o->x2, x3 = o->x3, x4 = o->x4, x5 = o->x5, x6 = o->x6;
    for (int q = 0; q < 7; q++) memset(o->LC[q], 0, sizeof(double) * nxy);
    const double *vr = o->src_mask;
#define VR(c, t) vr[(c) + nxy * (long)((t)-1)]
    for (long c = 0; c < nxy; c++) {
      if (x0 >= 0)
        o->LC[0][c] = VR(c, time1 - 3) / ((x0 - x1) * (x0 - x2) * (x0 - x3) * (x0 - x4) * (x0 - x5) * (x0 - x6));
      if (x1 >= 0)
        o->LC[1][c] = VR(c, time1 - 2) / ((x1 - x0) * (x1 - x2) * (x1 - x3) * (x1 - x4) * (x1 - x5) * (x1 - x6));
      if (x2 >= 0)
        o->LC[2][c] = VR(c, time1 - 1) / ((x2 - x0) * (x2 - x1) * (x2 - x3) * (x2 - x4) * (x2 - x5) * (x2 - x6));
      if (x3 >= 0)
        o->LC[3][c] = VR(c, time1) / ((x3 - x0) * (x3 - x1) * (x3 - x2) * (x3 - x4) * (x3 - x5) * (x3 - x6));
      o->LC[4][c] = VR(c, time1 + 1) / ((x4 - x0) * (x4 - x1) * (x4 - x2) * (x4 - x3) * (x4 - x5) * (x4 - x6));
      if (time1 < num_steps)
        o->LC[5][c] = VR(c, time1 + 2) / ((x5 - x0) * (x5 - x1) * (x5 - x2) * (x5 - x3) * (x5 - x4) * (x5 - x6));
      if (time1 < num_steps - 1)
        o->LC[6][c] = VR(c, time1 + 3) / ((x6 - x0) * (x6 - x1) * (x6 - x2) * (x6 - x3) * (x6 - x4) * (x6 - x5));
    }
#undef VR
    o->time_old = time1;
  }
  double x0 = o->x0, x1 = o->x1, x2 = o->x2, x3 = o->x3, x4 = o->x4, x5 = o->x5, x6 = o->x6;
  for (long c = 0; c < nxy; c++) {
    o->forcing[c] = o->LC[0][c] * (xt - x1) * (xt - x2) * (xt - x3) * (xt - x4) * (xt - x5) * (xt - x6) +
                    o->LC[1][c] * (xt - x0) * (xt - x2) * (xt - x3) * (xt - x4) * (xt - x5) * (xt - x6) +
                    o->LC[2][c] * (xt - x0) * (xt - x1) * (xt - x3) * (xt - x4) * (xt - x5) * (xt - x6) +
                    o->LC[3][c] * (xt - x0) * (xt - x1) * (xt - x2) * (xt - x4) * (xt - x5) * (xt - x6) +
                    o->LC[4][c] * (xt - x0) * (xt - x1) * (xt - x2) * (xt - x3) * (xt - x5) * (xt - x6) +
                    o->LC[5][c] * (xt - x0) * (xt - x1) * (xt - x2) * (xt - x3) * (xt - x4) * (xt - x6) +
                    o->LC[6][c] * (xt - x0) * (xt - x1) * (xt - x2) * (xt - x3) * (xt - x4) * (xt - x5);
  }
}

/* Standalone Lagrange evaluation for the tests: vr(nxy, nt) samples, time. */
int orc_lagrange_eval(const double *vr, long nxy, int nt, int time, double timestep,
                      double cad, int num_steps, double *out) {
  sparc_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  orc o;
  memset(&o, 0, sizeof(o));
  o.cfg = cfg;
  o.cfg.timestep = timestep;
  o.cfg.cadforcing = cad;
  o.cfg.num_steps = num_steps;
  o.nx = (int)nxy;
  o.ny = 1;
  o.src_mask = (double *)vr;
  o.forcing = out;
  for (int q = 0; q < 7; q++) o.LC[q] = zalloc(sizeof(double) * nxy);
  o.time = time;
  o.time_old = 0;
  (void)nt;
  lagrange_interp(&o);
  for (int q = 0; q < 7; q++) free(o.LC[q]);
  return SPARC_OK;
}

/* PML accounting for one derivative (displacement.f90:495-506):
 * RHSpsi = az*d + bzpml*psi; d = d + psi, on k <= npmlbot or k >= nz-npmltop+1 */
static void pml_account(orc *o, double *d, int which) {
  long nxy = (long)o->nx * o->ny;
  int pmlindex = 0;
  for (int k = 1; k <= o->nz; k++) {
    if (k <= o->cfg.npmlbot || k >= o->nz - o->cfg.npmltop + 1) {
      double *psi = o->psipml + which * o->NP + nxy * pmlindex;
      double *rhs = o->scrpml + which * o->NP + nxy * pmlindex;
      long off = nxy * (k - 1);
      for (long c = 0; c < nxy; c++) {
        rhs[c] = o->bg[SPARC_BG_AZ][off + c] * d[off + c] + o->bg[SPARC_BG_BZPML][off + c] * psi[c];
        d[off + c] = d[off + c] + psi[c];
      }
      pmlindex++;
    }
  }
}

/* damp_velocity, 3-D branch (damping.f90:125-160): RHSv -= IFFT2(rates*FFT2(v)) */
static void damp_field(orc *o, const double *v, double *rhs) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long nxy = (long)nx * ny;
  fftplan *px = plan_new(nx), *py = plan_new(ny);
#pragma omp parallel
  {
    cplx *pl = malloc(sizeof(cplx) * nxy), *b = malloc(sizeof(cplx) * (nx > ny ? nx : ny));
    cplx *s = malloc(sizeof(cplx) * (nx > ny ? nx : ny));
#pragma omp for schedule(static)
    for (int k = 0; k < nz; k++) {
      for (long c = 0; c < nxy; c++) pl[c] = v[c + nxy * k];
      for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) b[i] = pl[i + (long)nx * j];
        fft_rec(b, s, nx, 1, px, -1);
        for (int i = 0; i < nx; i++) pl[i + (long)nx * j] = s[i];
      }
      for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) b[j] = pl[i + (long)nx * j];
        fft_rec(b, s, ny, 1, py, -1);
        for (int j = 0; j < ny; j++) {
          int ii = (i <= nx / 2) ? i : nx - i;
          s[j] *= o->damping_rates[ii + (long)(nx / 2 + 1) * j];
        }
        fft_rec(s, b, ny, 1, py, +1);
        for (int j = 0; j < ny; j++) pl[i + (long)nx * j] = b[j];
      }
      for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) b[i] = pl[i + (long)nx * j];
        fft_rec(b, s, nx, 1, px, +1);
        for (int i = 0; i < nx; i++) rhs[i + (long)nx * j + nxy * k] = rhs[i + (long)nx * j + nxy * k] - creal(s[i]);
      }
    }
    free(pl);
    free(b);
    free(s);
  }
  plan_free(px);
  plan_free(py);
}

/* MP_QUIET_PML_3D_DISPL, option 4 (displacement.f90:482-658), kernel mode,
 * no horizontal PML. */
static void rhs_option4(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *xi_x = o->temp, *xi_y = o->temp + N, *xi_z = o->temp + 2 * N;
  double *v_x = o->temp + 3 * N, *v_y = o->temp + 4 * N, *v_z = o->temp + 5 * N;
  double **bg = o->bg;
  /* ddxyz(xi_x, dxixdx, xi_y, dxiydy, xi_z, dxizdz, bc=1) */
  orc_ddx(xi_x, o->dxdx, nx, ny, nz, o->cfg.stretchx);
  orc_ddy(xi_y, o->dydy, nx, ny, nz, o->cfg.stretchy);
  orc_ddz(xi_z, o->dzdz, nx, ny, nz, 1, o->prof[SPARC_PROF_STRETCH]);
  pml_account(o, o->dzdz, 0);
  for (long c = 0; c < N; c++) o->div[c] = o->dxdx[c] + o->dydy[c] + o->dzdz[c];
  for (long c = 0; c < N; c++) {
    o->p[c] = -bg[SPARC_BG_C2RHO0][c] * o->div[c] - xi_z[c] * bg[SPARC_BG_GRADP0_Z][c];
    o->rho[c] = -bg[SPARC_BG_RHO0][c] * o->div[c] - xi_z[c] * bg[SPARC_BG_GRADRHO0_Z][c];
  }
  /* ddxyz(p, gradp_x, p, gradp_y, p, gradp_z, bc=0) */
  orc_ddx(o->p, o->gpx, nx, ny, nz, o->cfg.stretchx);
  orc_ddy(o->p, o->gpy, nx, ny, nz, o->cfg.stretchy);
  orc_ddz(o->p, o->gpz, nx, ny, nz, 0, o->prof[SPARC_PROF_STRETCH]);
  pml_account(o, o->gpz, 1);
  double *RHSxi_x = o->scr, *RHSxi_y = o->scr + N, *RHSxi_z = o->scr + 2 * N;
  double *RHSv_x = o->scr + 3 * N, *RHSv_y = o->scr + 4 * N, *RHSv_z = o->scr + 5 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  for (long c = 0; c < N; c++) {
    RHSv_x[c] = -rhoinv[c] * o->gpx[c];
    RHSv_y[c] = -rhoinv[c] * o->gpy[c];
    RHSv_z[c] = -rhoinv[c] * o->gpz[c];
  }
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      RHSv_z[q] = RHSv_z[q] - o->rho[q] * g[k] * rhoinv[q];
    }
  for (long c = 0; c < N; c++) {
    RHSxi_x[c] = v_x[c];
    RHSxi_y[c] = v_y[c];
    RHSxi_z[c] = v_z[c];
  }
  /* kernel-mode source (displacement.f90:611-634) */
  if (o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until) {
    int kk = (o->cfg.mode == SPARC_ADJOINT) ? o->cfg.o_rad - 1 : o->cfg.e_rad - 1;
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * kk;
      RHSv_z[q] = RHSv_z[q] + o->forcing[c] / bg[SPARC_BG_RHO0][q] * o->cfg.delta_width;
    }
  }
  if (o->cfg.damp_waves) {
    damp_field(o, v_x, RHSv_x);
    damp_field(o, v_y, RHSv_y);
    damp_field(o, v_z, RHSv_z);
  }
  for (int f = 0; f < 6; f++)
    for (long c = 0; c < nxy; c++) {
      o->scr[f * N + nxy * (nz - 1) + c] = 0.0;
      o->scr[f * N + c] = 0.0;
    }
}

/* MP_QUIET_PML_3D, option 2 ((rho, v, p) form, pml.f90:12-76), no
 * horizontal PML.  Slots: rho = 0, v = 1..3, p = 4. */
static void rhs_option2(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *rho = o->temp, *v_x = o->temp + N, *v_y = o->temp + 2 * N, *v_z = o->temp + 3 * N;
  double *p = o->temp + 4 * N;
  double **bg = o->bg;
  /* ddxyz(v_x, dvxdx, v_y, dvydy, v_z, dvzdz, 1); ddxyz(p, gradp_x, p, gradp_y, p, gradp_z, 0) */
  orc_ddx(v_x, o->dxdx, nx, ny, nz, o->cfg.stretchx);
  orc_ddy(v_y, o->dydy, nx, ny, nz, o->cfg.stretchy);
  orc_ddz(v_z, o->dzdz, nx, ny, nz, 1, o->prof[SPARC_PROF_STRETCH]);
  orc_ddx(p, o->gpx, nx, ny, nz, o->cfg.stretchx);
  orc_ddy(p, o->gpy, nx, ny, nz, o->cfg.stretchy);
  orc_ddz(p, o->gpz, nx, ny, nz, 0, o->prof[SPARC_PROF_STRETCH]);
  /* PML_BOUNDARY_ACCOUNTING_3D: psip on gradp_z, psivz on dvzdz (pml.f90:222-228) */
  pml_account(o, o->gpz, 1);
  pml_account(o, o->dzdz, 0);
  for (long c = 0; c < N; c++) o->div[c] = o->dxdx[c] + o->dydy[c] + o->dzdz[c];
  double *RHScont = o->scr, *RHSv_x = o->scr + N, *RHSv_y = o->scr + 2 * N;
  double *RHSv_z = o->scr + 3 * N, *RHSp = o->scr + 4 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  const double *sdep = o->prof[SPARC_PROF_SOURCE_DEP];
  for (long c = 0; c < N; c++) {
    RHScont[c] = -bg[SPARC_BG_GRADRHO0_Z][c] * v_z[c] - bg[SPARC_BG_RHO0][c] * o->div[c];
    RHSv_x[c] = -rhoinv[c] * o->gpx[c];
    RHSv_y[c] = -rhoinv[c] * o->gpy[c];
  }
  int src = o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until;
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      RHSv_z[q] = -rhoinv[q] * (o->gpz[q] + rho[q] * g[k]);
      if (src && abs(k - (o->cfg.e_rad - 1)) <= 4) RHSv_z[q] = RHSv_z[q] + o->forcing[c] * sdep[k];
    }
  for (long c = 0; c < N; c++)
    RHSp[c] = -bg[SPARC_BG_C2RHO0][c] * o->div[c] - v_z[c] * bg[SPARC_BG_GRADP0_Z][c];
  /* zero planes 1 and nz of the five fields, then DAMP_VELOCITY on every plane */
  for (int f = 0; f < 5; f++)
    for (long c = 0; c < nxy; c++) {
      o->scr[f * N + nxy * (nz - 1) + c] = 0.0;
      o->scr[f * N + c] = 0.0;
    }
  if (o->cfg.damp_waves) {
    damp_field(o, v_x, RHSv_x);
    damp_field(o, v_y, RHSv_y);
    damp_field(o, v_z, RHSv_z);
  }
}

/* The PML of MP_MHD_PML_3D_DISPL: bottom layers only (displacement.f90:681-692) */
static void pml_account_bottom(orc *o, double *d, int which) {
  long nxy = (long)o->nx * o->ny;
  for (int k = 1; k <= o->cfg.npmlbot; k++) {
    double *psi = o->psipml + which * o->NP + nxy * (k - 1);
    double *rhs = o->scrpml + which * o->NP + nxy * (k - 1);
    long off = nxy * (k - 1);
    for (long c = 0; c < nxy; c++) {
      rhs[c] = o->bg[SPARC_BG_AZ][off + c] * d[off + c] + o->bg[SPARC_BG_BZPML][off + c] * psi[c];
      d[off + c] = d[off + c] + psi[c];
    }
  }
}

/* MP_MHD_PML_3D_DISPL, option 10 (displacement.f90:663-849), kernel mode,
 * sources at the planes e_rad/o_rad (erad_2d = e_rad everywhere).
 * ddxyz(f1,d1,f2,d2,f3,d3,bc): d1 = d/dx f1, d2 = d/dy f2, d3 = d/dz f3. */
static void rhs_option10(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *xi_x = o->temp, *xi_y = o->temp + N, *xi_z = o->temp + 2 * N;
  double *v_x = o->temp + 3 * N, *v_y = o->temp + 4 * N, *v_z = o->temp + 5 * N;
  double **bg = o->bg, **m = o->mw;
  const double *bx0 = bg[SPARC_BG_BX], *by0 = bg[SPARC_BG_BY], *bz0 = bg[SPARC_BG_BZ];
  const double *st = o->prof[SPARC_PROF_STRETCH];
  double sx = o->cfg.stretchx, sy = o->cfg.stretchy;
  double *f1 = m[0], *f2 = m[1], *f3 = m[2];
  /* flux = -(xi x B0) (:673-675) */
  for (long c = 0; c < N; c++) {
    f1[c] = -(xi_z[c] * by0[c] - xi_y[c] * bz0[c]);
    f2[c] = -(xi_x[c] * bz0[c] - xi_z[c] * bx0[c]);
    f3[c] = -(xi_y[c] * bx0[c] - xi_x[c] * by0[c]);
  }
  /* div xi with psivz, p, rho (:677-700) */
  orc_ddx(xi_x, o->dxdx, nx, ny, nz, sx);
  orc_ddy(xi_y, o->dydy, nx, ny, nz, sy);
  orc_ddz(xi_z, o->dzdz, nx, ny, nz, 1, st);
  pml_account_bottom(o, o->dzdz, 0);
  for (long c = 0; c < N; c++) {
    double div = o->dxdx[c] + o->dydy[c] + o->dzdz[c];
    o->p[c] = -bg[SPARC_BG_C2RHO0][c] * div - xi_z[c] * bg[SPARC_BG_GRADP0_Z][c] -
              xi_x[c] * bg[SPARC_BG_GRADP0_X][c] - xi_y[c] * bg[SPARC_BG_GRADP0_Y][c];
    o->rho[c] = -bg[SPARC_BG_RHO0][c] * div - xi_z[c] * bg[SPARC_BG_GRADRHO0_Z][c] -
                xi_x[c] * bg[SPARC_BG_GRADRHO0_X][c] - xi_y[c] * bg[SPARC_BG_GRADRHO0_Y][c];
  }
  /* grad p with psip (:702-717) */
  orc_ddx(o->p, o->gpx, nx, ny, nz, sx);
  orc_ddy(o->p, o->gpy, nx, ny, nz, sy);
  orc_ddz(o->p, o->gpz, nx, ny, nz, 0, st);
  pml_account_bottom(o, o->gpz, 1);
  /* b = curl(flux) (:719-753) */
  double *dxf3 = m[3], *dyf3 = m[4], *dzf2 = m[5], *dxf2 = m[6], *dyf1 = m[7], *dzf1 = m[8];
  orc_ddx(f3, dxf3, nx, ny, nz, sx);
  orc_ddy(f3, dyf3, nx, ny, nz, sy);
  orc_ddz(f2, dzf2, nx, ny, nz, 1, st);
  orc_ddx(f2, dxf2, nx, ny, nz, sx);
  orc_ddy(f1, dyf1, nx, ny, nz, sy);
  orc_ddz(f1, dzf1, nx, ny, nz, 1, st);
  pml_account_bottom(o, dzf2, 4); /* psiinductionbx */
  pml_account_bottom(o, dzf1, 5); /* psiinductionby */
  double *bx = f1, *by = f2, *bz = f3; /* fluxes are consumed */
  for (long c = 0; c < N; c++) {
    double b1 = dyf3[c] - dzf2[c], b2 = -dxf3[c] + dzf1[c], b3 = dxf2[c] - dyf1[c];
    bx[c] = b1;
    by[c] = b2;
    bz[c] = b3;
  }
  /* J = curl(b) (:755-788) */
  double *dxby = m[3], *dybz = m[4], *dzbx = m[5], *dxbz = m[6], *dybx = m[7], *dzby = m[8];
  orc_ddx(by, dxby, nx, ny, nz, sx);
  orc_ddy(bz, dybz, nx, ny, nz, sy);
  orc_ddz(bx, dzbx, nx, ny, nz, 1, st);
  orc_ddx(bz, dxbz, nx, ny, nz, sx);
  orc_ddy(bx, dybx, nx, ny, nz, sy);
  orc_ddz(by, dzby, nx, ny, nz, 1, st);
  pml_account_bottom(o, dzbx, 2); /* psidzbx */
  pml_account_bottom(o, dzby, 3); /* psidzby */
  double *RHSxi_x = o->scr, *RHSxi_y = o->scr + N, *RHSxi_z = o->scr + 2 * N;
  double *RHSv_x = o->scr + 3 * N, *RHSv_y = o->scr + 4 * N, *RHSv_z = o->scr + 5 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  const double *sponge = o->prof[SPARC_PROF_SPONGE];
  const double *jx0 = bg[SPARC_BG_CURLBX], *jy0 = bg[SPARC_BG_CURLBY], *jz0 = bg[SPARC_BG_CURLBZ];
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      double jx = dybz[q] - dzby[q], jy = -dxbz[q] + dzbx[q], jz = dxby[q] - dybx[q];
      /* cross(J, B0) and cross(J0, b) (:790-801) */
      double l1 = jy * bz0[q] - jz * by0[q], l2 = jz * bx0[q] - jx * bz0[q];
      double l3 = jx * by0[q] - jy * bx0[q];
      double m1 = jy0[q] * bz[q] - jz0[q] * by[q], m2 = jz0[q] * bx[q] - jx0[q] * bz[q];
      double m3 = jx0[q] * by[q] - jy0[q] * bx[q];
      double rx = rhoinv[q] * (-o->gpx[q] + l1) + rhoinv[q] * m1;
      double ry = rhoinv[q] * (-o->gpy[q] + l2) + rhoinv[q] * m2;
      double rz = rhoinv[q] * (-o->gpz[q] + l3) + rhoinv[q] * m3;
      if (k + 1 >= nz - 35) { /* top sponge (:803-807) */
        rx = rx - sponge[k] * v_x[q];
        ry = ry - sponge[k] * v_y[q];
        rz = rz - sponge[k] * v_z[q];
      }
      rz = rz - rhoinv[q] * o->rho[q] * g[k];
      RHSv_x[q] = rx;
      RHSv_y[q] = ry;
      RHSv_z[q] = rz;
      RHSxi_x[q] = v_x[q];
      RHSxi_y[q] = v_y[q];
      RHSxi_z[q] = v_z[q];
    }
  if (o->cfg.damp_waves) {
    damp_field(o, v_x, RHSv_x);
    damp_field(o, v_y, RHSv_y);
    damp_field(o, v_z, RHSv_z);
  }
  /* kernel-mode source at orad_2d / erad_2d (:813-836) */
  if (o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until) {
    int kk = (o->cfg.mode == SPARC_ADJOINT) ? o->cfg.o_rad - 1 : o->cfg.e_rad - 1;
    for (long c = 0; c < nxy; c++) {
      int kc = o->src_planes ? o->src_planes[c] : kk;
      double dw = o->src_planes ? o->src_dw[c] : o->cfg.delta_width;
      long q = c + nxy * kc;
      RHSv_z[q] = RHSv_z[q] + o->forcing[c] * dw / bg[SPARC_BG_RHO0][q];
    }
  }
  for (int f = 0; f < 6; f++)
    for (long c = 0; c < nxy; c++) {
      o->scr[f * N + nxy * (nz - 1) + c] = 0.0;
      o->scr[f * N + c] = 0.0;
    }
}

/* MP_MHD_PML_3D, option 8 ((rho, v, p, b) form, pml.f90:80-208), no
 * horizontal PML.  Slots: rho 0, v 1..3, p 4, b 5..7. */
static void rhs_option8(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *rho = o->temp, *vx = o->temp + N, *vy = o->temp + 2 * N, *vz = o->temp + 3 * N;
  double *p = o->temp + 4 * N, *bx = o->temp + 5 * N, *by = o->temp + 6 * N, *bz = o->temp + 7 * N;
  double **bg = o->bg, **m = o->mw;
  const double *st = o->prof[SPARC_PROF_STRETCH];
  double sx = o->cfg.stretchx, sy = o->cfg.stretchy;
  const double *b0x = bg[SPARC_BG_BX], *b0y = bg[SPARC_BG_BY], *b0z = bg[SPARC_BG_BZ];
  double *f1 = m[0], *f2 = m[1], *f3 = m[2];
  /* flux = v x B0 (:89-91: flux1 = -(v_z boy - v_y boz), ...) */
  for (long c = 0; c < N; c++) {
    f1[c] = -(vz[c] * b0y[c] - vy[c] * b0z[c]);
    f2[c] = -(vx[c] * b0z[c] - vz[c] * b0x[c]);
    f3[c] = -(vy[c] * b0x[c] - vx[c] * b0y[c]);
  }
  /* the six ddxyz calls (:104-115) */
  double *dxf3 = m[3], *dzf1 = m[4], *dyf3 = m[5], *dzf2 = m[6], *dxf2 = m[7], *dybz = m[8];
  double *dzbx = m[9], *dxby = m[10], *dybx = m[11], *dxbz = o->div, *dyf1 = o->p, *dzby = o->rho;
  orc_ddx(vx, o->dxdx, nx, ny, nz, sx);
  orc_ddy(vy, o->dydy, nx, ny, nz, sy);
  orc_ddz(vz, o->dzdz, nx, ny, nz, 1, st);
  orc_ddx(f3, dxf3, nx, ny, nz, sx);
  orc_ddy(p, o->gpy, nx, ny, nz, sy);
  orc_ddz(f1, dzf1, nx, ny, nz, 1, st);
  orc_ddx(p, o->gpx, nx, ny, nz, sx);
  orc_ddy(f3, dyf3, nx, ny, nz, sy);
  orc_ddz(f2, dzf2, nx, ny, nz, 1, st);
  orc_ddx(f2, dxf2, nx, ny, nz, sx);
  orc_ddy(bz, dybz, nx, ny, nz, sy);
  orc_ddz(bx, dzbx, nx, ny, nz, 1, st);
  orc_ddx(by, dxby, nx, ny, nz, sx);
  orc_ddy(bx, dybx, nx, ny, nz, sy);
  orc_ddz(p, o->gpz, nx, ny, nz, 1, st);
  orc_ddx(bz, dxbz, nx, ny, nz, sx);
  orc_ddy(f1, dyf1, nx, ny, nz, sy);
  orc_ddz(by, dzby, nx, ny, nz, 1, st);
  /* PML_BOUNDARY_ACCOUNTING_3D with the magnetic terms (pml.f90:213-280) */
  pml_account(o, o->gpz, 1);
  pml_account(o, o->dzdz, 0);
  pml_account(o, dzf2, 4);
  pml_account(o, dzf1, 5);
  pml_account(o, dzbx, 2);
  pml_account(o, dzby, 3);
  double *RHScont = o->scr, *RHSv_x = o->scr + N, *RHSv_y = o->scr + 2 * N, *RHSv_z = o->scr + 3 * N;
  double *RHSp = o->scr + 4 * N, *RHSb_x = o->scr + 5 * N, *RHSb_y = o->scr + 6 * N;
  double *RHSb_z = o->scr + 7 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  const double *sdep = o->prof[SPARC_PROF_SOURCE_DEP];
  const double *j0x = bg[SPARC_BG_CURLBX], *j0y = bg[SPARC_BG_CURLBY], *j0z = bg[SPARC_BG_CURLBZ];
  int src = o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until;
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      double rbx = dyf3[q] - dzf2[q], rby = -dxf3[q] + dzf1[q], rbz = dxf2[q] - dyf1[q];
      double div = o->dxdx[q] + o->dydy[q] + o->dzdz[q];
      double jx = dybz[q] - dzby[q], jy = -dxbz[q] + dzbx[q], jz = dxby[q] - dybx[q];
      double rc = -bg[SPARC_BG_GRADRHO0_X][q] * vx[q] - bg[SPARC_BG_GRADRHO0_Y][q] * vy[q] -
                  bg[SPARC_BG_GRADRHO0_Z][q] * vz[q] - bg[SPARC_BG_RHO0][q] * div;
      double rp = -bg[SPARC_BG_C2RHO0][q] * div - vx[q] * bg[SPARC_BG_GRADP0_X][q] -
                  vy[q] * bg[SPARC_BG_GRADP0_Y][q] - vz[q] * bg[SPARC_BG_GRADP0_Z][q];
      double l1 = jy * b0z[q] - jz * b0y[q], l2 = b0x[q] * jz - jx * b0z[q];
      double l3 = jx * b0y[q] - jy * b0x[q];
      double rx = -rhoinv[q] * (o->gpx[q] - l1), ry = -rhoinv[q] * (o->gpy[q] - l2);
      double rz = -rhoinv[q] * (o->gpz[q] - l3);
      double m1 = j0y[q] * bz[q] - j0z[q] * by[q], m2 = bx[q] * j0z[q] - j0x[q] * bz[q];
      double m3 = j0x[q] * by[q] - j0y[q] * bx[q];
      rx = rx + rhoinv[q] * m1;
      ry = ry + rhoinv[q] * m2;
      rz = rz + rhoinv[q] * m3;
      rz = rz - rhoinv[q] * rho[q] * g[k];
      if (src && abs(k - (o->cfg.e_rad - 1)) <= 4) rz = rz + o->forcing[c] * sdep[k];
      RHScont[q] = rc;
      RHSv_x[q] = rx;
      RHSv_y[q] = ry;
      RHSv_z[q] = rz;
      RHSp[q] = rp;
      RHSb_x[q] = rbx;
      RHSb_y[q] = rby;
      RHSb_z[q] = rbz;
    }
  for (int f = 0; f < 8; f++)
    for (long c = 0; c < nxy; c++) {
      o->scr[f * N + nxy * (nz - 1) + c] = 0.0;
      o->scr[f * N + c] = 0.0;
    }
  if (o->cfg.damp_waves) {
    damp_field(o, vx, RHSv_x);
    damp_field(o, vy, RHSv_y);
    damp_field(o, vz, RHSv_z);
  }
}

/* MP_MHD_PML_2D_DISPL, option 18 (physics2d.f90:724-943), ny = 1, kernel mode,
 * sources at the planes e_rad/o_rad (erad_2d uniform). */
static void rhs_option18(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *xi_x = o->temp, *xi_z = o->temp + 2 * N;
  double *v_x = o->temp + 3 * N, *v_z = o->temp + 5 * N;
  double **bg = o->bg, **m = o->mw;
  const double *st = o->prof[SPARC_PROF_STRETCH], *unstretch = o->prof[SPARC_PROF_UNSTRETCH];
  double sx = o->cfg.stretchx;
  /* ddx(xi_x); dxizdz(nz) preset; ddz(xi_z, 3); psivz (:735-752) */
  orc_ddx(xi_x, o->dxdx, nx, ny, nz, sx);
  for (long c = 0; c < nxy; c++)
    o->dzdz[c + nxy * (nz - 1)] = -o->dxdx[c + nxy * (nz - 1)] * unstretch[nz - 1];
  orc_ddz(xi_z, o->dzdz, nx, ny, nz, 3, st);
  pml_account(o, o->dzdz, 0);
  for (long c = 0; c < N; c++) {
    double div = o->dxdx[c] + o->dzdz[c];
    o->p[c] = -bg[SPARC_BG_C2RHO0][c] * div - xi_z[c] * bg[SPARC_BG_GRADP0_Z][c] -
              xi_x[c] * bg[SPARC_BG_GRADP0_X][c];
    o->rho[c] = -bg[SPARC_BG_RHO0][c] * div - xi_z[c] * bg[SPARC_BG_GRADRHO0_Z][c] -
                xi_x[c] * bg[SPARC_BG_GRADRHO0_X][c];
  }
  /* ddx(p); ddz(p, 2) with gradp_z(:,:,1) the caller's (zero) value; psip (:775-791) */
  orc_ddx(o->p, o->gpx, nx, ny, nz, sx);
  for (long c = 0; c < nxy; c++) o->gpz[c] = 0.0;
  orc_ddz(o->p, o->gpz, nx, ny, nz, 2, st);
  pml_account(o, o->gpz, 1);
  /* flux2, bx = -(ddz(flux2, 0) + psiinductionbx), bz = ddx(flux2) (:815-845) */
  double *f2 = m[0], *bx = m[1], *bz = m[2], *dzbx = m[3], *cby = m[4];
  for (long c = 0; c < N; c++)
    f2[c] = -(xi_x[c] * bg[SPARC_BG_BZ][c] - xi_z[c] * bg[SPARC_BG_BX][c]);
  orc_ddz(f2, bx, nx, ny, nz, 0, st);
  orc_ddx(f2, bz, nx, ny, nz, sx);
  pml_account(o, bx, 4); /* psiinductionbx */
  for (long c = 0; c < N; c++) bx[c] = -bx[c];
  /* dzbx = ddz(bx, 0) + psidzbx, curlby = -ddx(bz) + dzbx (:852-873) */
  orc_ddz(bx, dzbx, nx, ny, nz, 0, st);
  orc_ddx(bz, cby, nx, ny, nz, sx);
  pml_account(o, dzbx, 2); /* psidzbx */
  double *RHSxi_x = o->scr, *RHSxi_z = o->scr + 2 * N;
  double *RHSv_x = o->scr + 3 * N, *RHSv_z = o->scr + 5 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      double curlby = -cby[q] + dzbx[q];
      double fl1 = curlby * bg[SPARC_BG_BZ][q], fl3 = -curlby * bg[SPARC_BG_BX][q];
      double rx = rhoinv[q] * (-o->gpx[q] + fl1);
      double rz = rhoinv[q] * (-o->gpz[q] + fl3);
      fl1 = bg[SPARC_BG_CURLBY][q] * bz[q];
      fl3 = -bg[SPARC_BG_CURLBY][q] * bx[q];
      rx = rx + rhoinv[q] * fl1;
      rz = rz + rhoinv[q] * fl3;
      rz = rz - rhoinv[q] * o->rho[q] * g[k];
      RHSv_x[q] = rx;
      RHSv_z[q] = rz;
      RHSxi_x[q] = v_x[q];
      RHSxi_z[q] = v_z[q];
    }
  /* DAMP_VELOCITY, 2-D branch: v_x and v_z (damping.f90:162-178) */
  if (o->cfg.damp_waves) {
    damp_field(o, v_x, RHSv_x);
    damp_field(o, v_z, RHSv_z);
  }
  if (o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until) {
    int kk = (o->cfg.mode == SPARC_ADJOINT) ? o->cfg.o_rad - 1 : o->cfg.e_rad - 1;
    for (long c = 0; c < nxy; c++) {
      int kc = o->src_planes ? o->src_planes[c] : kk;
      double dw = o->src_planes ? o->src_dw[c] : o->cfg.delta_width;
      long q = c + nxy * kc;
      RHSv_z[q] = RHSv_z[q] + o->forcing[c] * dw / bg[SPARC_BG_RHO0][q];
    }
  }
  /* only plane 1 is zeroed (:930-933) */
  for (int f = 0; f < 6; f++)
    for (long c = 0; c < nxy; c++) o->scr[f * N + c] = 0.0;
}

/* MP_QUIET_PML_2D_DISPL, option 17 (physics2d.f90:948-1105), ny = 1. */
static void rhs_option17(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N, nxy = (long)nx * ny;
  double *xi_x = o->temp, *xi_z = o->temp + 2 * N;
  double *v_x = o->temp + 3 * N, *v_z = o->temp + 5 * N;
  double **bg = o->bg;
  const double *unstretch = o->prof[SPARC_PROF_UNSTRETCH];
  orc_ddx(xi_x, o->dxdx, nx, ny, nz, o->cfg.stretchx);
  for (long c = 0; c < nxy; c++)
    o->dzdz[c + nxy * (nz - 1)] = -o->dxdx[c + nxy * (nz - 1)] * unstretch[nz - 1];
  orc_ddz(xi_z, o->dzdz, nx, ny, nz, 3, o->prof[SPARC_PROF_STRETCH]);
  pml_account(o, o->dzdz, 0);
  for (long c = 0; c < N; c++) o->div[c] = o->dxdx[c] + o->dzdz[c];
  for (long c = 0; c < N; c++)
    o->p[c] = -bg[SPARC_BG_C2RHO0][c] * o->div[c] - bg[SPARC_BG_GRADP0_Z][c] * xi_z[c];
  orc_ddx(o->p, o->gpx, nx, ny, nz, o->cfg.stretchx);
  /* ibc=2: gradp_z(:,:,1) is whatever the array holds (physics2d.f90:1018);
   * the persistent gpz buffer starts at zero and stays zero (psip(1) starts at
   * zero and its RHS is az*0 + bz*0). */
  orc_ddz(o->p, o->gpz, nx, ny, nz, 2, o->prof[SPARC_PROF_STRETCH]);
  pml_account(o, o->gpz, 1);
  double *RHSxi_x = o->scr, *RHSxi_z = o->scr + 2 * N;
  double *RHSv_x = o->scr + 3 * N, *RHSv_z = o->scr + 5 * N;
  const double *rhoinv = bg[SPARC_BG_RHOINV], *g = o->prof[SPARC_PROF_G];
  for (long c = 0; c < N; c++) RHSv_x[c] = -rhoinv[c] * o->gpx[c];
  for (long c = 0; c < N; c++)
    o->rho[c] = -bg[SPARC_BG_RHO0][c] * o->div[c] - bg[SPARC_BG_GRADRHO0_Z][c] * xi_z[c];
  for (int k = 0; k < nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      RHSv_z[q] = -rhoinv[q] * (o->gpz[q] + o->rho[q] * g[k]);
    }
  for (long c = 0; c < N; c++) {
    RHSxi_x[c] = v_x[c];
    RHSxi_z[c] = v_z[c];
  }
  if (o->src_kind != SPARC_SRC_NONE && o->time <= o->src_active_until) {
    int kk = (o->cfg.mode == SPARC_ADJOINT) ? o->cfg.o_rad - 1 : o->cfg.e_rad - 1;
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * kk;
      RHSv_z[q] = RHSv_z[q] + o->forcing[c] * o->cfg.delta_width / bg[SPARC_BG_RHO0][q];
    }
  }
  for (long c = 0; c < nxy; c++) {
    RHSv_x[c] = 0.;
    RHSv_z[c] = 0.;
    RHSxi_x[c] = 0.;
    RHSxi_z[c] = 0.;
  }
}

static const double FILT_C[6] = {0.75390625000000, 0.41015625000000, -0.23437500000000,
                                 0.08789062500000, -0.01953125000000, 0.00195312500000};

/* 11-point z filter of one field over planes kst..nz-1 (1-based), with the
 * 2-point averages at 2..5 and nz-4..nz-1 (physics.f90:760-788,
 * physics2d.f90:527-560, kernels.f90:845-881). */
static void zfilter_field(double *f, int nx, int ny, int nz, int kst, double *tmp) {
  long nxy = (long)nx * ny;
#define F_(k) (f + nxy * (long)((k)-1))
#define T_(k) (tmp + nxy * (long)((k)-1))
  for (long c = 0; c < nxy; c++) {
    T_(2)[c] = 0.5 * (F_(3)[c] + F_(1)[c]);
    T_(3)[c] = 0.5 * (F_(4)[c] + F_(2)[c]);
    T_(4)[c] = 0.5 * (F_(5)[c] + F_(3)[c]);
    T_(5)[c] = 0.5 * (F_(6)[c] + F_(4)[c]);
  }
#pragma omp parallel for schedule(static)
  for (int k = 6; k <= nz - 5; k++)
    for (long c = 0; c < nxy; c++)
      T_(k)[c] = FILT_C[0] * F_(k)[c] + 0.5 * FILT_C[1] * (F_(k - 1)[c] + F_(k + 1)[c]) +
                 0.5 * FILT_C[2] * (F_(k - 2)[c] + F_(k + 2)[c]) +
                 0.5 * FILT_C[3] * (F_(k - 3)[c] + F_(k + 3)[c]) +
                 0.5 * FILT_C[4] * (F_(k - 4)[c] + F_(k + 4)[c]) +
                 0.5 * FILT_C[5] * (F_(k - 5)[c] + F_(k + 5)[c]);
  for (long c = 0; c < nxy; c++) {
    T_(nz - 4)[c] = 0.5 * (F_(nz - 3)[c] + F_(nz - 5)[c]);
    T_(nz - 3)[c] = 0.5 * (F_(nz - 2)[c] + F_(nz - 4)[c]);
    T_(nz - 2)[c] = 0.5 * (F_(nz - 1)[c] + F_(nz - 3)[c]);
    T_(nz - 1)[c] = 0.5 * (F_(nz)[c] + F_(nz - 2)[c]);
  }
  for (int k = kst; k <= nz - 1; k++) memcpy(F_(k), T_(k), sizeof(double) * nxy);
#undef F_
#undef T_
}

void orc_zfilter(double *f, int nx, int ny, int nz, int kst) {
  double *tmp = zalloc(sizeof(double) * (long)nx * ny * nz);
  zfilter_field(f, nx, ny, nz, kst, tmp);
  free(tmp);
}

/* FILTER_VARS_3D (physics.f90:741-833) / FILTER_VARS_2D (physics2d.f90:514-588) */
static void filter_vars(orc *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  long N = o->N;
  int is2d = (o->cfg.option == SPARC_OPT_QUIET_PML_2D_DISPL ||
              o->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL);
  for (int l = 0; l < o->nf; l++) zfilter_field(o->a + l * N, nx, ny, nz, is2d ? 6 : 2, o->term);
  int ff = o->cfg.freq_filtering;
  if (ff <= 0) return;
  if (!is2d) {
    if (o->time % ff == 0) {
      for (int l = 0; l < 5; l++) lowpass_axis(o->a + l * N, nx, ny, nz, 0, o->decay_x);
      for (int l = 0; l < 5; l++) lowpass_axis(o->a + l * N, nx, ny, nz, 1, o->decay_y);
    }
  } else {
    if ((o->time + 1) % ff == 0)
      for (int l = 0; l < 6; l++) lowpass_axis(o->a + l * N, nx, ny, nz, 0, o->decay_x);
  }
}

/* STEP (step.f90:43-128) */
static void step(orc *o) {
  long N6 = (long)o->nf * o->N, NP2 = (long)o->npv * o->NP;
  int npml = o->nzpml > 0;
  for (int s = 0; s < 5; s++) {
    if (s == 0) {
      memcpy(o->temp, o->a, sizeof(double) * N6);
      if (npml) memcpy(o->psipml, o->pmlvars, sizeof(double) * NP2);
    } else {
      double b = o->betas[s];
      for (long c = 0; c < N6; c++) o->temp[c] = o->a[c] + b * o->scr[c];
      if (npml)
        for (long c = 0; c < NP2; c++) o->psipml[c] = o->pmlvars[c] + b * o->scrpml[c];
    }
    if (o->cfg.option == SPARC_OPT_QUIET_PML_3D_DISPL)
      rhs_option4(o);
    else if (o->cfg.option == SPARC_OPT_QUIET_PML_3D)
      rhs_option2(o);
    else if (o->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL)
      rhs_option10(o);
    else if (o->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL)
      rhs_option18(o);
    else if (o->cfg.option == SPARC_OPT_MHD_PML_3D)
      rhs_option8(o);
    else
      rhs_option17(o);
  }
  double dt = o->cfg.deltat;
  for (long c = 0; c < N6; c++) o->a[c] = o->a[c] + dt * o->scr[c];
  if (npml)
    for (long c = 0; c < NP2; c++) o->pmlvars[c] = o->pmlvars[c] + dt * o->scrpml[c];
  filter_vars(o);
}

/* TIMESTEPPING inner sequence (driver.f90:412-424) */
int orc_step(orc *o, int t0, int nsteps) {
  long nxy = (long)o->nx * o->ny;
  for (int i = t0; i < t0 + nsteps; i++) {
    o->time = i;
    if (o->src_kind == SPARC_SRC_LAGRANGE && i <= o->src_active_until) lagrange_interp(o);
    if (o->src_kind == SPARC_SRC_SERIES) {
      double s = (i < o->src_nt) ? o->src_series[i] : 0.0;
      for (long c = 0; c < nxy; c++) o->forcing[c] = o->src_mask[c] * s;
    }
    step(o);
  }
  return SPARC_OK;
}

static double *arr_ptr(orc *o, int arr, int field, long *n) {
  if (arr == SPARC_ARR_A && field >= 0 && field < o->nf) {
    *n = o->N;
    return o->a + field * o->N;
  }
  if (arr == SPARC_ARR_SCR && field >= 0 && field < o->nf) {
    *n = o->N;
    return o->scr + field * o->N;
  }
  if (arr == SPARC_ARR_PMLVARS && field >= 0 && field < o->npv) {
    *n = o->NP;
    return o->pmlvars + field * o->NP;
  }
  if (arr == SPARC_ARR_SCRPML && field >= 0 && field < o->npv) {
    *n = o->NP;
    return o->scrpml + field * o->NP;
  }
  return NULL;
}

int orc_get_state(orc *o, int arr, int field, double *h) {
  long n;
  double *p = arr_ptr(o, arr, field, &n);
  if (!p) return SPARC_EINVAL;
  memcpy(h, p, sizeof(double) * n);
  return SPARC_OK;
}

int orc_put_state(orc *o, int arr, int field, const double *h) {
  long n;
  double *p = arr_ptr(o, arr, field, &n);
  if (!p) return SPARC_EINVAL;
  memcpy(p, h, sizeof(double) * n);
  return SPARC_OK;
}

/* norm (all_modules.f90:176-199) */
double orc_norm(orc *o, int field) {
  double s = 0.0;
  const double *f = o->a + field * o->N;
  for (long c = 0; c < o->N; c++) s = s + pow(f[c], 2.0);
  return sqrt(s / ((double)o->nx * (double)o->ny * (double)o->nz));
}

/* ======================================================================
 * Kernels (kernels.f90:197-313): snapshot pairs accumulated in FP64.
 * ====================================================================== */
/* fields per snapshot: xi, acc (+ v with the flow kernels) */
static int snap_fields(const orc *o) { return o->cfg.flow_kernels ? 9 : 6; }

int orc_snapshots_alloc(orc *o, int nslots) {
  free(o->snap);
  o->nslots = nslots;
  o->snap = zalloc(sizeof(double) * snap_fields(o) * o->N * (size_t)nslots);
  return SPARC_OK;
}

/* WRITE_OUT_PARTIAL_STATE (all_modules.f90:760-776): xi = a(1..3), acc = scr(4..6),
 * v = a(4..6) */
int orc_snapshot_store(orc *o, int slot) {
  if (slot < 0 || slot >= o->nslots) return SPARC_EINVAL;
  double *s = o->snap + (size_t)slot * snap_fields(o) * o->N;
  memcpy(s, o->a, sizeof(double) * 3 * o->N);
  memcpy(s + 3 * o->N, o->scr + 3 * o->N, sizeof(double) * 3 * o->N);
  if (o->cfg.flow_kernels) memcpy(s + 6 * o->N, o->a + 3 * o->N, sizeof(double) * 3 * o->N);
  return SPARC_OK;
}

int orc_snapshot_get(orc *o, int slot, int f, double *h) {
  if (slot < 0 || slot >= o->nslots || f < 0 || f >= snap_fields(o)) return SPARC_EINVAL;
  memcpy(h, o->snap + ((size_t)slot * snap_fields(o) + f) * o->N, sizeof(double) * o->N);
  return SPARC_OK;
}

int orc_snapshot_put(orc *o, int slot, int f, const double *h) {
  if (slot < 0 || slot >= o->nslots || f < 0 || f >= snap_fields(o)) return SPARC_EINVAL;
  memcpy(o->snap + ((size_t)slot * snap_fields(o) + f) * o->N, h, sizeof(double) * o->N);
  return SPARC_OK;
}

int orc_kernel_begin(orc *o, const double *grav, int nz) {
  if (nz != o->nz) return SPARC_EINVAL;
  for (int q = 0; q < SPARC_KERN_COUNT; q++) memset(o->kern[q], 0, sizeof(double) * o->N);
  memcpy(o->grav, grav, sizeof(double) * nz);
  return SPARC_OK;
}

/* ddzkern (kernels.f90:598-611): dbyd2 * stretchkern, then FILTER_VAR_Z */
static void ddzkern(orc *o, const double *in, double *out, int bc, double *tmp) {
  orc_ddz(in, out, o->nx, o->ny, o->nz, bc, o->prof[SPARC_PROF_STRETCH]);
  zfilter_field(out, o->nx, o->ny, o->nz, 2, tmp);
}

/* COMPUTE_DIV (kernels.f90:798-816) */
static void compute_div(orc *o, const double *fx, const double *fy, const double *fz,
                        double *dive, double *t1, double *t2) {
  long N = o->N;
  orc_ddx(fx, dive, o->nx, o->ny, o->nz, o->cfg.stretchx);
  if (o->ny > 4) {
    orc_ddy(fy, t1, o->nx, o->ny, o->nz, o->cfg.stretchy);
    for (long c = 0; c < N; c++) dive[c] = dive[c] + t1[c];
  }
  ddzkern(o, fz, t1, 1, t2);
  for (long c = 0; c < N; c++) dive[c] = dive[c] + t1[c];
}

/* CROSS_KERN (kernels.f90:757-766): c = a x b */
static void cross3(const double *ax, const double *ay, const double *az, const double *bx,
                   const double *by, const double *bz, double *cx, double *cy, double *cz,
                   long N) {
  for (long c = 0; c < N; c++) {
    double x = ay[c] * bz[c] - az[c] * by[c], y = bx[c] * az[c] - ax[c] * bz[c];
    double z = ax[c] * by[c] - ay[c] * bx[c];
    cx[c] = x;
    cy[c] = y;
    cz[c] = z;
  }
}

/* CURL_KERN (kernels.f90:770-793); its bcz is an unassigned module integer
 * (initialize.f90:145), zero in static storage, so the z derivatives take
 * ibc 0 (and FILTER_VAR_Z through ddzkern) */
static void curl_kern(orc *o, const double *fx, const double *fy, const double *fz, double *cx,
                      double *cy, double *cz, double *t, double *t2) {
  long N = o->N;
  orc_ddx(fy, cz, o->nx, o->ny, o->nz, o->cfg.stretchx);
  orc_ddy(fx, t, o->nx, o->ny, o->nz, o->cfg.stretchy);
  for (long c = 0; c < N; c++) cz[c] = cz[c] - t[c];
  orc_ddy(fz, cx, o->nx, o->ny, o->nz, o->cfg.stretchy);
  ddzkern(o, fy, t, 0, t2);
  for (long c = 0; c < N; c++) cx[c] = cx[c] - t[c];
  ddzkern(o, fx, cy, 0, t2);
  orc_ddx(fz, t, o->nx, o->ny, o->nz, o->cfg.stretchx);
  for (long c = 0; c < N; c++) cy[c] = cy[c] - t[c];
}

/* The MAGNETIC_KERNELS sums (kernels.f90:336-430) for one snapshot pair */
static void magnetic_kernels(orc *o, const double *fx, const double *fy, const double *fz,
                             const double *ax, const double *ay, const double *az) {
  long N = o->N;
  double **bg = o->bg;
  const double *b0x = bg[SPARC_BG_BX], *b0y = bg[SPARC_BG_BY], *b0z = bg[SPARC_BG_BZ];
  const double *j0x = bg[SPARC_BG_CURLBX], *j0y = bg[SPARC_BG_CURLBY], *j0z = bg[SPARC_BG_CURLBZ];
  double *kx = o->kern[SPARC_KERN_BX], *ky = o->kern[SPARC_KERN_BY], *kz = o->kern[SPARC_KERN_BZ];
  double **m = o->mw;
  double *tx = m[0], *ty = m[1], *tz = m[2], *fbx = m[3], *fby = m[4], *fbz = m[5];
  double *ux = m[6], *uy = m[7], *uz = m[8], *wx = m[9], *wy = m[10], *wz = m[11];
  double *s1 = o->term, *s2 = o->div, *d = o->p;
  /* v_a x curl(curl(xi x B0)) */
  cross3(fx, fy, fz, b0x, b0y, b0z, tx, ty, tz, N);
  curl_kern(o, tx, ty, tz, fbx, fby, fbz, s1, s2);
  curl_kern(o, fbx, fby, fbz, ux, uy, uz, s1, s2);
  cross3(ax, ay, az, ux, uy, uz, wx, wy, wz, N);
  for (long c = 0; c < N; c++) {
    kx[c] += wx[c];
    ky[c] += wy[c];
    kz[c] += wz[c];
  }
  /* curl(fb x v_a) */
  cross3(fbx, fby, fbz, ax, ay, az, tx, ty, tz, N);
  curl_kern(o, tx, ty, tz, wx, wy, wz, s1, s2);
  for (long c = 0; c < N; c++) {
    kx[c] += wx[c];
    ky[c] += wy[c];
    kz[c] += wz[c];
  }
  /* xi x curl(curl(v_a x B0)) */
  cross3(ax, ay, az, b0x, b0y, b0z, tx, ty, tz, N);
  curl_kern(o, tx, ty, tz, fbx, fby, fbz, s1, s2);
  curl_kern(o, fbx, fby, fbz, ux, uy, uz, s1, s2);
  cross3(fx, fy, fz, ux, uy, uz, wx, wy, wz, N);
  for (long c = 0; c < N; c++) {
    kx[c] += wx[c];
    ky[c] += wy[c];
    kz[c] += wz[c];
  }
  /* curl(v_a x J0) x xi */
  cross3(ax, ay, az, j0x, j0y, j0z, wx, wy, wz, N);
  curl_kern(o, wx, wy, wz, tx, ty, tz, s1, s2);
  cross3(tx, ty, tz, fx, fy, fz, wx, wy, wz, N);
  for (long c = 0; c < N; c++) {
    kx[c] += wx[c];
    ky[c] += wy[c];
    kz[c] += wz[c];
  }
  /* (xi x J0) div(v_a) */
  compute_div(o, ax, ay, az, d, s1, s2);
  cross3(fx, fy, fz, j0x, j0y, j0z, tx, ty, tz, N);
  for (long c = 0; c < N; c++) {
    kx[c] += tx[c] * d[c];
    ky[c] += ty[c] * d[c];
    kz[c] += tz[c] * d[c];
  }
  /* curl(B0 x (xi div(v_a))) */
  for (long c = 0; c < N; c++) {
    tx[c] = fx[c] * d[c];
    ty[c] = fy[c] * d[c];
    tz[c] = fz[c] * d[c];
  }
  cross3(b0x, b0y, b0z, tx, ty, tz, wx, wy, wz, N);
  curl_kern(o, wx, wy, wz, tx, ty, tz, s1, s2);
  for (long c = 0; c < N; c++) {
    kx[c] += tx[c];
    ky[c] += ty[c];
    kz[c] += tz[c];
  }
}

int orc_kernel_accumulate(orc *o, int slot) {
  if (slot < 0 || slot >= o->nslots) return SPARC_EINVAL;
  long N = o->N, nxy = (long)o->nx * o->ny;
  const double *s = o->snap + (size_t)slot * snap_fields(o) * N;
  const double *fxx = s, *fxy = s + N, *fxz = s + 2 * N;
  const double *fax = s + 3 * N, *fay = s + 4 * N, *faz = s + 5 * N;
  const double *avx = o->a + 3 * N, *avy = o->a + 4 * N, *avz = o->a + 5 * N;
  double *temp1 = zalloc(sizeof(double) * N), *temp2 = zalloc(sizeof(double) * N);
  double *temp3 = zalloc(sizeof(double) * N), *t1 = zalloc(sizeof(double) * N);
  double *t2 = zalloc(sizeof(double) * N), *d31 = zalloc(sizeof(double) * N);
  double *d32 = zalloc(sizeof(double) * N), *d33 = zalloc(sizeof(double) * N);
  double *hess = o->kern[SPARC_KERN_HESSIAN], *kc2 = o->kern[SPARC_KERN_C2];
  double *krho = o->kern[SPARC_KERN_RHO];
  const double *c2 = o->bg[SPARC_BG_C2];
  for (long c = 0; c < N; c++)
    hess[c] = hess[c] + faz[c] * faz[c] + fax[c] * fax[c] + fay[c] * fay[c];
  compute_div(o, fxx, fxy, fxz, temp2, t1, t2);
  compute_div(o, avx, avy, avz, temp3, t1, t2);
  for (long c = 0; c < N; c++) {
    temp1[c] = temp2[c] * temp3[c];
    kc2[c] = kc2[c] + temp1[c];
  }
  for (long c = 0; c < N; c++)
    krho[c] = krho[c] + temp1[c] * c2[c] + faz[c] * avz[c] + fax[c] * avx[c] + fay[c] * avy[c];
  for (int k = 0; k < o->nz; k++)
    for (long c = 0; c < nxy; c++) temp2[c + nxy * k] = avz[c + nxy * k] * o->grav[k];
  orc_ddx(temp2, d31, o->nx, o->ny, o->nz, o->cfg.stretchx);
  orc_ddy(temp2, d32, o->nx, o->ny, o->nz, o->cfg.stretchy);
  ddzkern(o, temp2, d33, 1, t2);
  for (int k = 0; k < o->nz; k++)
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      krho[q] = krho[q] - o->grav[k] * fxz[q] * temp3[q] +
                (fxx[q] * d31[q] + fxy[q] * d32[q] + fxz[q] * d33[q]);
    }
  if (o->cfg.option == SPARC_OPT_MHD_PML_3D_DISPL || o->cfg.option == SPARC_OPT_MHD_PML_2D_DISPL)
    magnetic_kernels(o, fxx, fxy, fxz, avx, avy, avz);
  if (o->cfg.flow_kernels) {
    /* COMPUTE_DEL (kernels.f90:821-841): deriv(i,j) = d_i f_vel_j; kernelv(i) +=
     * sum_j deriv(i,j) * a_vel_j (kernels.f90:317-332); no y part when ny <= 4 */
    const double *av[3] = {avx, avy, avz};
    double *kv[3] = {o->kern[SPARC_KERN_VX], o->kern[SPARC_KERN_VY], o->kern[SPARC_KERN_VZ]};
    for (int j = 0; j < 3; j++) {
      const double *fv = s + (size_t)(6 + j) * N;
      orc_ddx(fv, d31, o->nx, o->ny, o->nz, o->cfg.stretchx);
      if (o->ny > 4) orc_ddy(fv, d32, o->nx, o->ny, o->nz, o->cfg.stretchy);
      else memset(d32, 0, sizeof(double) * N);
      ddzkern(o, fv, d33, 1, t2);
      for (long c = 0; c < N; c++) {
        kv[0][c] = kv[0][c] + d31[c] * av[j][c];
        kv[1][c] = kv[1][c] + d32[c] * av[j][c];
        kv[2][c] = kv[2][c] + d33[c] * av[j][c];
      }
    }
  }
  free(temp1);
  free(temp2);
  free(temp3);
  free(t1);
  free(t2);
  free(d31);
  free(d32);
  free(d33);
  return SPARC_OK;
}

int orc_kernel_get(orc *o, int which, double *h) {
  if (which < 0 || which >= SPARC_KERN_COUNT) return SPARC_EINVAL;
  if (which == SPARC_KERN_PSI) {
    /* kernel_vectorpsi (kernels.f90:508-509): the y component of
     * CURL_KERN(K_Bx, 0, K_Bz) = ddzkern(K_Bx, ibc 0) - ddx(K_Bz) */
    long N = o->N;
    double *t = zalloc(sizeof(double) * N), *t2 = zalloc(sizeof(double) * N);
    ddzkern(o, o->kern[SPARC_KERN_BX], o->kern[SPARC_KERN_PSI], 0, t2);
    orc_ddx(o->kern[SPARC_KERN_BZ], t, o->nx, o->ny, o->nz, o->cfg.stretchx);
    for (long c = 0; c < N; c++) o->kern[SPARC_KERN_PSI][c] -= t[c];
    free(t);
    free(t2);
  }
  memcpy(h, o->kern[which], sizeof(double) * o->N);
  return SPARC_OK;
}

/* sets a kernel sum directly, so tests can pin what is formed from the sums */
int orc_kernel_put(orc *o, int which, const double *h) {
  if (which < 0 || which >= SPARC_KERN_COUNT) return SPARC_EINVAL;
  memcpy(o->kern[which], h, sizeof(double) * o->N);
  return SPARC_OK;
}

/* ======================================================================
 * Travel times: COMPUTE_TT (driver.f90:921-960) with inverse (964-1042)
 * ====================================================================== */
static void inverse3(double a[3][3], double c[3][3]) {
  int n = 3;
  double L[3][3] = {{0}}, U[3][3] = {{0}}, b[3] = {0}, d[3], x[3];
  for (int k = 0; k < n - 1; k++)
    for (int i = k + 1; i < n; i++) {
      double coeff = a[i][k] / a[k][k];
      L[i][k] = coeff;
      for (int j = k + 1; j < n; j++) a[i][j] = a[i][j] - coeff * a[k][j];
    }
  for (int i = 0; i < n; i++) L[i][i] = 1.0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i <= j; i++) U[i][j] = a[i][j];
  for (int k = 0; k < n; k++) {
    b[k] = 1.0;
    d[0] = b[0];
    for (int i = 1; i < n; i++) {
      d[i] = b[i];
      for (int j = 0; j < i; j++) d[i] = d[i] - L[i][j] * d[j];
    }
    x[n - 1] = d[n - 1] / U[n - 1][n - 1];
    for (int i = n - 2; i >= 0; i--) {
      x[i] = d[i];
      for (int j = n - 1; j > i; j--) x[i] = x[i] - U[i][j] * x[j];
      x[i] = x[i] / U[i][i];
    }
    for (int i = 0; i < n; i++) c[i][k] = x[i];
    b[k] = 0.0;
  }
}

void orc_inverse3(const double *a_in, double *c_out) {
  double a[3][3], c[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = a_in[i * 3 + j];
  inverse3(a, c);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c_out[i * 3 + j] = c[i][j];
}

double orc_compute_tt(const double *u0, const double *u, int nt, double dt) {
  int nc = 2 * nt - 1;
  double *cc = malloc(sizeof(double) * nc), *t = malloc(sizeof(double) * nc);
#define CC(i) cc[(i) + nt - 1]
#define TT(i) t[(i) + nt - 1]
  for (int i = -nt + 1; i <= 0; i++) {
    TT(i) = i * dt;
    int l1 = -i + 1, l2 = nt;
    double s = 0;
    for (int q = 0; q <= l2 - l1; q++) s = s + u0[l1 - 1 + q] * u[q];
    CC(i) = s * dt;
  }
  for (int i = 1; i <= nt - 1; i++) {
    TT(i) = i * dt;
    int l1 = i + 1, l2 = nt - i;
    double s = 0;
    for (int q = 0; q < l2; q++) s = s + u0[q] * u[l1 - 1 + q];
    CC(i) = s * dt;
  }
  /* maxloc returns the first maximum (1-based) */
  int best = 0;
  for (int q = 1; q < nc; q++)
    if (cc[q] > cc[best]) best = q;
  int loc = (best + 1) - nt;
  double times[3] = {TT(loc - 1), TT(loc), TT(loc + 1)};
  double mat[3][3], inv[3][3];
  for (int r = 0; r < 3; r++) {
    mat[r][0] = 1;
    mat[r][1] = times[r];
    mat[r][2] = pow(times[r], 2.0);
  }
  inverse3(mat, inv);
  double p1 = inv[1][0] * CC(loc - 1) + inv[1][1] * CC(loc) + inv[1][2] * CC(loc + 1);
  double p2 = inv[2][0] * CC(loc - 1) + inv[2][1] * CC(loc) + inv[2][2] * CC(loc + 1);
#undef CC
#undef TT
  free(cc);
  free(t);
  return -p1 * 0.5 / p2;
}

/* ======================================================================
 * Set-up restatement: the derived background of mp_Initialize_RHS
 * (physics.f90:204-303) / mp_Initialize_RHS_2d (physics2d.f90:235-342).
 * Inputs: z(nz), c2, rho0, p0 (N each, nondimensional), g(nz).
 * Outputs: rhoinv, c2rho0, gradp0_z, gradrho0_z (PML-decayed), az, bzpml (N),
 * g (PML-decayed, nz), unstretch/stretch (nz).
 * ====================================================================== */
int orc_init_rhs(int nx, int ny, int nz, int is2d, int npmlbot, int npmltop, double diml,
                 double dimc, const double *z, const double *c2, const double *rho0,
                 const double *p0, double *g, double *rhoinv, double *c2rho0,
                 double *gradp0_z, double *gradrho0_z, double *az, double *bzpml,
                 double *unstretch, double *stretch) {
  long nxy = (long)nx * ny, N = nxy * nz;
  /* initialize.f90:464-468 */
  orc_dbyd2(unstretch, z, 1, nz, 1);
  for (int k = 0; k < nz; k++) stretch[k] = 1.0 / unstretch[k];
  for (long c = 0; c < N; c++) {
    rhoinv[c] = 1.0 / rho0[c];
    c2rho0[c] = c2[c] * rho0[c];
  }
  orc_ddz(rho0, gradrho0_z, nx, ny, nz, 1, stretch);
  orc_ddz(p0, gradp0_z, nx, ny, nz, 1, stretch);
  double *tempin = malloc(sizeof(double) * N), *dampz = calloc(N, sizeof(double));
  double *alphaz = malloc(sizeof(double) * nz);
  for (long c = 0; c < N; c++) tempin[c] = pow(c2[c], 0.5);
  double pd = PI_R * SP(0.005) * diml / dimc;
  for (int k = 0; k < nz; k++) alphaz[k] = pd;
  /* the 3-D bottom profile carries an extra factor 2 (physics.f90:268 vs
   * physics2d.f90:319); the 2-D top uses c at nz-5 (physics2d.f90:329) */
  double botfac = is2d ? 3. * 3. : 2. * 3. * 3.;
  long topk = is2d ? nz - 6 : nz - 1;
  for (int k1 = 1; k1 <= nz; k1++) {
    int k = k1 - 1;
    if (k1 < npmlbot + 1) {
      double L = z[npmlbot - 1] - z[0];
      alphaz[k] = PI_R * SP(0.005) * diml / dimc * (z[k] - z[0]) / L;
      for (long c = 0; c < nxy; c++) {
        long q = c + nxy * k;
        dampz[q] = botfac * tempin[c] / L * pow((z[npmlbot - 1] - z[k]) / L, 2.0);
        double fd = alphaz[k] / (dampz[q] + alphaz[k]);
        gradp0_z[q] = gradp0_z[q] * fd;
        gradrho0_z[q] = gradrho0_z[q] * fd;
      }
      g[k] = g[k] * (alphaz[k] / (dampz[nxy * k] + alphaz[k]));
    }
    if (k1 > nz - npmltop) {
      double L = z[nz - 1] - z[nz - npmltop];
      alphaz[k] = PI_R * SP(0.005) * diml / dimc * (z[nz - 1] - z[k]) / L;
      for (long c = 0; c < nxy; c++) {
        long q = c + nxy * k;
        dampz[q] = 3. * 3. * tempin[c + nxy * topk] / L * pow((z[nz - npmltop] - z[k]) / L, 2.0);
        double fd = alphaz[k] / (dampz[q] + alphaz[k]);
        gradp0_z[q] = gradp0_z[q] * fd;
        gradrho0_z[q] = gradrho0_z[q] * fd;
      }
      g[k] = g[k] * (alphaz[k] / (dampz[nxy * k] + alphaz[k]));
    }
    for (long c = 0; c < nxy; c++) {
      long q = c + nxy * k;
      bzpml[q] = -alphaz[k] - dampz[q];
      az[q] = -dampz[q];
    }
  }
  free(tempin);
  free(dampz);
  free(alphaz);
  return SPARC_OK;
}

/* damp_velocity for one field, exposed for the tests (damping.f90:125-160) */
int orc_damp_field(orc *o, const double *v, double *rhs) {
  damp_field(o, v, rhs);
  return SPARC_OK;
}
