/*
 * oracle.c -- CPU FP64 restatement of exafmm-beta's Laplace FMM path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
 * operator definitions; each function names the file:line it follows.
 * Compiled without -ffast-math and with -ffp-contract=off so the tree,
 * the multipole acceptance test and the operators round like the reference's
 * scalar x86 build.
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;
#define EPS 1e-16 /* include/types.h:19 */

struct orc_kernel {
  int P, NTERM;
  double eps2;
  double *prefactor; /* 4P^2 */
  double *Anm;       /* 4P^2 */
  cplx *Cnm;         /* P^4 */
};

static int odd_or_even(int n) { return (n & 1) ? -1 : 1; } /* laplace.h:24-26 */

/* i^e for integer e (the reference evaluates std::pow(I, real(e))). */
static cplx ipow(int e) {
  int r = ((e % 4) + 4) % 4;
  return r == 0 ? 1.0 : r == 1 ? I : r == 2 ? -1.0 : -I;
}

orc_kernel *orc_kernel_create(int P, double eps2) {
  orc_kernel *k = (orc_kernel *)calloc(1, sizeof(orc_kernel));
  k->P = P;
  k->NTERM = P * (P + 1) / 2;
  k->eps2 = eps2;
  k->prefactor = (double *)calloc(4 * P * P, sizeof(double));
  k->Anm = (double *)calloc(4 * P * P, sizeof(double));
  k->Cnm = (cplx *)calloc((size_t)P * P * P * P, sizeof(cplx));
  /* laplace.h:121-136 */
  for (int n = 0; n < 2 * P; n++) {
    for (int m = -n; m <= n; m++) {
      int nm = n * n + n + m;
      int nabsm = abs(m);
      double fnmm = EPS;
      for (int i = 1; i <= n - m; i++) fnmm *= i;
      double fnpm = EPS;
      for (int i = 1; i <= n + m; i++) fnpm *= i;
      double fnma = 1.0;
      for (int i = 1; i <= n - nabsm; i++) fnma *= i;
      double fnpa = 1.0;
      for (int i = 1; i <= n + nabsm; i++) fnpa *= i;
      k->prefactor[nm] = sqrt(fnma / fnpa);
      k->Anm[nm] = odd_or_even(n) / sqrt(fnmm * fnpm);
    }
  }
  /* laplace.h:137-147 */
  for (int j = 0, jk = 0, jknm = 0; j < P; j++) {
    for (int kk = -j; kk <= j; kk++, jk++) {
      for (int n = 0, nm = 0; n < P; n++) {
        for (int m = -n; m <= n; m++, nm++, jknm++) {
          int jnkm = (j + n) * (j + n) + j + n + m - kk;
          k->Cnm[jknm] = ipow(abs(kk - m) - abs(kk) - abs(m)) *
                         (double)(odd_or_even(j) * k->Anm[nm] * k->Anm[jk] / k->Anm[jnkm]) * EPS;
        }
      }
    }
  }
  return k;
}

void orc_kernel_destroy(orc_kernel *k) {
  if (!k) return;
  free(k->prefactor);
  free(k->Anm);
  free(k->Cnm);
  free(k);
}

int orc_kernel_nterm(const orc_kernel *k) { return k->NTERM; }

static double norm3(const double *d) { return d[0] * d[0] + d[1] * d[1] + d[2] * d[2]; }

/* laplace.h:29-33 */
static void cart2sph(const double *dX, double *r, double *theta, double *phi) {
  *r = sqrt(norm3(dX));
  *theta = *r == 0 ? 0 : acos(dX[2] / *r);
  *phi = atan2(dX[1], dX[0]);
}

/* laplace.h:36-45 */
static void sph2cart(double r, double theta, double phi, const double *sp, double *c) {
  c[0] = sin(theta) * cos(phi) * sp[0] + cos(theta) * cos(phi) / r * sp[1] -
         sin(phi) / r / sin(theta) * sp[2];
  c[1] = sin(theta) * sin(phi) * sp[0] + cos(theta) * sin(phi) / r * sp[1] +
         cos(phi) / r / sin(theta) * sp[2];
  c[2] = cos(theta) * sp[0] - sin(theta) / r * sp[1];
}

/* laplace.h:48-80 */
static void eval_multipole(const orc_kernel *k, double rho, double alpha, double beta,
                           cplx *Ynm, cplx *YnmTheta) {
  int P = k->P;
  double x = cos(alpha), y = sin(alpha);
  double fact = 1, pn = 1, rhom = 1;
  for (int m = 0; m < P; m++) {
    cplx eim = cexp(I * (double)(m * beta));
    double p = pn;
    int npn = m * m + 2 * m, nmn = m * m;
    Ynm[npn] = rhom * p * k->prefactor[npn] * eim;
    Ynm[nmn] = conj(Ynm[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    YnmTheta[npn] = rhom * (p - (m + 1) * x * p1) / y * k->prefactor[npn] * eim;
    rhom *= rho;
    double rhon = rhom;
    for (int n = m + 1; n < P; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      Ynm[npm] = rhon * p * k->prefactor[npm] * eim;
      Ynm[nmm] = conj(Ynm[npm]);
      double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      YnmTheta[npm] = rhon * ((n - m + 1) * p - (n + 1) * x * p1) / y * k->prefactor[npm] * eim;
      rhon *= rho;
    }
    pn = -pn * fact * y;
    fact += 2;
  }
}

/* laplace.h:83-113 */
static void eval_local(const orc_kernel *k, double rho, double alpha, double beta, cplx *Ynm2) {
  int P = k->P;
  double x = cos(alpha), y = sin(alpha);
  double fact = 1, pn = 1, rhom = 1.0 / rho;
  for (int m = 0; m < 2 * P; m++) {
    cplx eim = cexp(I * (double)(m * beta));
    double p = pn;
    int npn = m * m + 2 * m, nmn = m * m;
    Ynm2[npn] = rhom * p * k->prefactor[npn] * eim;
    Ynm2[nmn] = conj(Ynm2[npn]);
    double p1 = p;
    p = x * (2 * m + 1) * p1;
    rhom /= rho;
    double rhon = rhom;
    for (int n = m + 1; n < 2 * P; n++) {
      int npm = n * n + n + m, nmm = n * n + n - m;
      Ynm2[npm] = rhon * p * k->prefactor[npm] * eim;
      Ynm2[nmm] = conj(Ynm2[npm]);
      double p2 = p1;
      p1 = p;
      p = (x * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
      rhon /= rho;
    }
    pn = -pn * fact * y;
    fact += 2;
  }
}

/* laplace.h:212-234 (scalar path; the SIMD path differs only in rsqrt rounding) */
void orc_p2p(const orc_kernel *k, int ni, const double *xi, double *trgi, int nj,
             const double *xj, const double *qj, const double *xper) {
  double xp[3] = {0, 0, 0};
  if (xper) memcpy(xp, xper, sizeof xp);
  for (int i = 0; i < ni; i++) {
    double pot = 0, ax = 0, ay = 0, az = 0;
    for (int j = 0; j < nj; j++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = xi[3 * i + d] - xj[3 * j + d] - xp[d];
      double R2 = norm3(dX) + k->eps2;
      if (R2 != 0) {
        double invR2 = 1.0 / R2;
        double invR = qj[j] * sqrt(invR2);
        double s = invR2 * invR;
        pot += invR;
        ax += dX[0] * s;
        ay += dX[1] * s;
        az += dX[2] * s;
      }
    }
    trgi[4 * i + 0] += pot;
    trgi[4 * i + 1] -= ax;
    trgi[4 * i + 2] -= ay;
    trgi[4 * i + 3] -= az;
  }
}

/* laplace.h:237-252 */
void orc_p2m(const orc_kernel *k, const double *Xc, int nb, const double *x, const double *q,
             double *Mout) {
  int P = k->P;
  cplx *M = (cplx *)Mout;
  cplx *Ynm = (cplx *)malloc(sizeof(cplx) * P * P), *Yt = (cplx *)malloc(sizeof(cplx) * P * P);
  for (int b = 0; b < nb; b++) {
    double dX[3], rho, alpha, beta;
    for (int d = 0; d < 3; d++) dX[d] = x[3 * b + d] - Xc[d];
    cart2sph(dX, &rho, &alpha, &beta);
    eval_multipole(k, rho, alpha, -beta, Ynm, Yt);
    for (int n = 0; n < P; n++)
      for (int m = 0; m <= n; m++) M[n * (n + 1) / 2 + m] += q[b] * Ynm[n * n + n + m];
  }
  free(Ynm);
  free(Yt);
}

/* laplace.h:254-290, one child */
void orc_m2m(const orc_kernel *k, const double *Xp, const double *Xc, const double *Mcin,
             double *Mpout) {
  int P = k->P;
  const cplx *Mc = (const cplx *)Mcin;
  cplx *Mp = (cplx *)Mpout;
  cplx *Ynm = (cplx *)malloc(sizeof(cplx) * P * P), *Yt = (cplx *)malloc(sizeof(cplx) * P * P);
  double dX[3], rho, alpha, beta;
  for (int d = 0; d < 3; d++) dX[d] = Xp[d] - Xc[d];
  cart2sph(dX, &rho, &alpha, &beta);
  eval_multipole(k, rho, alpha, -beta, Ynm, Yt);
  for (int j = 0; j < P; j++) {
    for (int kk = 0; kk <= j; kk++) {
      int jk = j * j + j + kk, jks = j * (j + 1) / 2 + kk;
      cplx M = 0;
      for (int n = 0; n <= j; n++) {
        int mmax = kk - 1 < n ? kk - 1 : n;
        for (int m = -n; m <= mmax; m++) {
          if (j - n >= kk - m) {
            int jnkm = (j - n) * (j - n) + j - n + kk - m;
            int jnkms = (j - n) * (j - n + 1) / 2 + kk - m;
            int nm = n * n + n + m;
            M += Mc[jnkms] * ipow(m - abs(m)) * Ynm[nm] *
                 (double)(odd_or_even(n) * k->Anm[nm] * k->Anm[jnkm] / k->Anm[jk]);
          }
        }
        for (int m = kk; m <= n; m++) {
          if (j - n >= m - kk) {
            int jnkm = (j - n) * (j - n) + j - n + kk - m;
            int jnkms = (j - n) * (j - n + 1) / 2 - kk + m;
            int nm = n * n + n + m;
            M += conj(Mc[jnkms]) * Ynm[nm] *
                 (double)(odd_or_even(kk + n + m) * k->Anm[nm] * k->Anm[jnkm] / k->Anm[jk]);
          }
        }
      }
      Mp[jks] += M * EPS;
    }
  }
  free(Ynm);
  free(Yt);
}

/* laplace.h:292-322 */
void orc_m2l(const orc_kernel *k, const double *Xi, const double *Xj, const double *xper,
             const double *Mjin, double *Liout) {
  int P = k->P;
  const cplx *Mj = (const cplx *)Mjin;
  cplx *Li = (cplx *)Liout;
  cplx *Ynm2 = (cplx *)malloc(sizeof(cplx) * 4 * P * P);
  double dX[3], rho, alpha, beta;
  for (int d = 0; d < 3; d++) dX[d] = Xi[d] - Xj[d] - (xper ? xper[d] : 0.0);
  cart2sph(dX, &rho, &alpha, &beta);
  eval_local(k, rho, alpha, beta, Ynm2);
  for (int j = 0; j < P; j++) {
    for (int kk = 0; kk <= j; kk++) {
      int jk = j * j + j + kk, jks = j * (j + 1) / 2 + kk;
      cplx L = 0;
      for (int n = 0; n < P; n++) {
        for (int m = -n; m < 0; m++) {
          int nm = n * n + n + m, nms = n * (n + 1) / 2 - m;
          int jknm = jk * P * P + nm, jnkm = (j + n) * (j + n) + j + n + m - kk;
          L += conj(Mj[nms]) * k->Cnm[jknm] * Ynm2[jnkm];
        }
        for (int m = 0; m <= n; m++) {
          int nm = n * n + n + m, nms = n * (n + 1) / 2 + m;
          int jknm = jk * P * P + nm, jnkm = (j + n) * (j + n) + j + n + m - kk;
          L += Mj[nms] * k->Cnm[jknm] * Ynm2[jnkm];
        }
      }
      Li[jks] += L;
    }
  }
  free(Ynm2);
}

/* laplace.h:324-357 */
void orc_l2l(const orc_kernel *k, const double *Xc, const double *Xp, const double *Lpin,
             double *Lcout) {
  int P = k->P;
  const cplx *Lp = (const cplx *)Lpin;
  cplx *Lc = (cplx *)Lcout;
  cplx *Ynm = (cplx *)malloc(sizeof(cplx) * P * P), *Yt = (cplx *)malloc(sizeof(cplx) * P * P);
  double dX[3], rho, alpha, beta;
  for (int d = 0; d < 3; d++) dX[d] = Xc[d] - Xp[d];
  cart2sph(dX, &rho, &alpha, &beta);
  eval_multipole(k, rho, alpha, beta, Ynm, Yt);
  for (int j = 0; j < P; j++) {
    for (int kk = 0; kk <= j; kk++) {
      int jk = j * j + j + kk, jks = j * (j + 1) / 2 + kk;
      cplx L = 0;
      for (int n = j; n < P; n++) {
        for (int m = j + kk - n; m < 0; m++) {
          int jnkm = (n - j) * (n - j) + n - j + m - kk;
          int nm = n * n + n - m, nms = n * (n + 1) / 2 - m;
          L += conj(Lp[nms]) * Ynm[jnkm] *
               (double)(odd_or_even(kk) * k->Anm[jnkm] * k->Anm[jk] / k->Anm[nm]);
        }
        for (int m = 0; m <= n; m++) {
          if (n - j >= abs(m - kk)) {
            int jnkm = (n - j) * (n - j) + n - j + m - kk;
            int nm = n * n + n + m, nms = n * (n + 1) / 2 + m;
            L += Lp[nms] * ipow(m - kk - abs(m - kk)) * Ynm[jnkm] * k->Anm[jnkm] * k->Anm[jk] /
                 k->Anm[nm];
          }
        }
      }
      Lc[jks] += L * EPS;
    }
  }
  free(Ynm);
  free(Yt);
}

/* laplace.h:359-388 */
void orc_l2p(const orc_kernel *k, const double *Xc, const double *Lin, int nb, const double *x,
             double *trg) {
  int P = k->P;
  const cplx *L = (const cplx *)Lin;
  cplx *Ynm = (cplx *)malloc(sizeof(cplx) * P * P), *Yt = (cplx *)malloc(sizeof(cplx) * P * P);
  for (int b = 0; b < nb; b++) {
    double dX[3], r, theta, phi, sp[3] = {0, 0, 0}, c[3];
    for (int d = 0; d < 3; d++) dX[d] = x[3 * b + d] - Xc[d] + EPS;
    cart2sph(dX, &r, &theta, &phi);
    eval_multipole(k, r, theta, phi, Ynm, Yt);
    double *t = trg + 4 * b;
    for (int n = 0; n < P; n++) {
      int nm = n * n + n, nms = n * (n + 1) / 2;
      t[0] += creal(L[nms] * Ynm[nm]);
      sp[0] += creal(L[nms] * Ynm[nm]) / r * n;
      sp[1] += creal(L[nms] * Yt[nm]);
      for (int m = 1; m <= n; m++) {
        nm = n * n + n + m;
        nms = n * (n + 1) / 2 + m;
        t[0] += 2 * creal(L[nms] * Ynm[nm]);
        sp[0] += 2 * creal(L[nms] * Ynm[nm]) / r * n;
        sp[1] += 2 * creal(L[nms] * Yt[nm]);
        sp[2] += 2 * creal(L[nms] * Ynm[nm] * I) * m;
      }
    }
    sph2cart(r, theta, phi, sp, c);
    t[1] += c[0];
    t[2] += c[1];
    t[3] += c[2];
  }
  free(Ynm);
  free(Yt);
}

/* bound_box.h:11-25 */
void orc_bounds(int n, const double *x, double *xmin, double *xmax) {
  if (n == 0) {
    for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = 0;
    return;
  }
  for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = x[d];
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      double a = xmin[d] - 1e-5, b = xmax[d] + 1e-5;
      xmin[d] = x[3 * i + d] < a ? x[3 * i + d] : a;
      xmax[d] = x[3 * i + d] > b ? x[3 * i + d] : b;
    }
  }
}

/* bound_box.h:28-37: getBounds(bodies, bounds) widens given bounds */
void orc_bounds_update(int n, const double *x, double *xmin, double *xmax) {
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      double a = xmin[d] - 1e-5, b = xmax[d] + 1e-5;
      xmin[d] = x[3 * i + d] < a ? x[3 * i + d] : a;
      xmax[d] = x[3 * i + d] > b ? x[3 * i + d] : b;
    }
  }
}

/* ---------------- tree (build_tree.h) ---------------- */
typedef struct onode {
  int ibody, nbody, nnode;
  struct onode *child[8];
  double X[3];
} onode;

struct orc_tree {
  int n, ncell, depth;
  int *perm;
  int *iparent, *ichild, *nchild, *ibody, *nbody;
  double *X, *R;
  uint64_t *icell;
};

typedef struct {
  const double *x; /* input-order coordinates */
  int ncrit;
} build_ctx;

/* countBodies/moveBodies/buildNodes: build_tree.h:29-51, 77-108.  `bodies`
 * and `buffer` hold input indices; the ping-pong and the leaf copy-back follow
 * the reference so the final order lands in the first array. */
static onode *build_nodes(build_ctx *c, int *bodies, int *buffer, int begin, int end,
                          const double *X, double R0, int level, int direction) {
  if (begin == end) return NULL;
  onode *nd = (onode *)calloc(1, sizeof(onode));
  nd->ibody = begin;
  nd->nbody = end - begin;
  nd->nnode = 1;
  memcpy(nd->X, X, sizeof(double) * 3);
  if (end - begin <= c->ncrit) {
    if (direction)
      for (int i = begin; i < end; i++) buffer[i] = bodies[i];
    return nd;
  }
  int nb[8] = {0}, off[8];
  for (int i = begin; i < end; i++) {
    const double *xb = c->x + 3 * (size_t)bodies[i];
    int oct = (xb[0] > X[0]) + ((xb[1] > X[1]) << 1) + ((xb[2] > X[2]) << 2);
    nb[oct]++;
  }
  for (int i = 0, o = begin; i < 8; i++) {
    off[i] = o;
    o += nb[i];
  }
  int start[8];
  memcpy(start, off, sizeof start);
  for (int i = begin; i < end; i++) {
    const double *xb = c->x + 3 * (size_t)bodies[i];
    int oct = (xb[0] > X[0]) + ((xb[1] > X[1]) << 1) + ((xb[2] > X[2]) << 2);
    buffer[off[oct]++] = bodies[i];
  }
  for (int i = 0; i < 8; i++) {
    double Xc[3];
    double r = R0 / (1 << (level + 1));
    for (int d = 0; d < 3; d++) {
      Xc[d] = X[d];
      Xc[d] += r * (((i & 1 << d) >> d) * 2 - 1);
    }
    nd->child[i] = build_nodes(c, buffer, bodies, start[i], start[i] + nb[i], Xc, R0, level + 1,
                               !direction);
  }
  for (int i = 0; i < 8; i++)
    if (nd->child[i]) nd->nnode += nd->child[i]->nnode;
  return nd;
}

/* getKey: build_tree.h:111-120 (with a 64-bit level offset) */
static uint64_t get_key(const double *X, const double *Xmin, double diameter, int level) {
  int iX[3];
  for (int d = 0; d < 3; d++) iX[d] = (int)((X[d] - Xmin[d]) / diameter);
  uint64_t index = (((uint64_t)1 << (3 * level)) - 1) / 7;
  for (int l = 0; l < level; l++) {
    for (int d = 0; d < 3; d++) index += (uint64_t)(iX[d] & 1) << (3 * l + d);
    for (int d = 0; d < 3; d++) iX[d] >>= 1;
  }
  return index;
}

/* nodes2cells: build_tree.h:123-165 */
static void nodes2cells(orc_tree *t, onode *nd, int C, int CN, const double *X0, double R0,
                        int level, int iparent) {
  t->iparent[C] = iparent;
  t->R[C] = R0 / (1 << level);
  memcpy(t->X + 3 * C, nd->X, sizeof(double) * 3);
  t->nbody[C] = nd->nbody;
  t->ibody[C] = nd->ibody;
  double Xmin[3] = {X0[0] - R0, X0[1] - R0, X0[2] - R0};
  t->icell[C] = get_key(t->X + 3 * C, Xmin, 2 * t->R[C], level);
  if (nd->nnode == 1) {
    t->ichild[C] = 0;
    t->nchild[C] = 0;
    if (level > t->depth) t->depth = level;
  } else {
    int nchild = 0, octants[8];
    for (int i = 0; i < 8; i++)
      if (nd->child[i]) octants[nchild++] = i;
    int Ci = CN;
    t->ichild[C] = Ci;
    t->nchild[C] = nchild;
    CN += nchild;
    for (int i = 0; i < nchild; i++) {
      onode *ch = nd->child[octants[i]];
      nodes2cells(t, ch, Ci, CN, X0, R0, level + 1, C);
      Ci++;
      CN += ch->nnode - 1;
    }
    for (int i = 0; i < nchild; i++) free(nd->child[octants[i]]);
    if (level + 1 > t->depth) t->depth = level + 1;
  }
}

orc_tree *orc_tree_build(int n, const double *x, int ncrit, const double *xmin,
                         const double *xmax) {
  orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
  t->n = n;
  t->perm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  if (n == 0) return t;
  /* bounds2box: build_tree.h:168-180 */
  double X[3], R = 0;
  for (int d = 0; d < 3; d++) X[d] = (xmax[d] + xmin[d]) / 2;
  for (int d = 0; d < 3; d++) {
    double a = X[d] - xmin[d];
    R = a > R ? a : R;
    double b = xmax[d] - X[d];
    R = b > R ? b : R;
  }
  R *= 1.00001;
  int *buffer = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) t->perm[i] = i;
  build_ctx c = {x, ncrit};
  onode *root = build_nodes(&c, t->perm, buffer, 0, n, X, R, 0, 0);
  free(buffer);
  t->ncell = root->nnode;
  int nc = t->ncell;
  t->iparent = (int *)malloc(sizeof(int) * nc);
  t->ichild = (int *)malloc(sizeof(int) * nc);
  t->nchild = (int *)malloc(sizeof(int) * nc);
  t->ibody = (int *)malloc(sizeof(int) * nc);
  t->nbody = (int *)malloc(sizeof(int) * nc);
  t->X = (double *)malloc(sizeof(double) * 3 * nc);
  t->R = (double *)malloc(sizeof(double) * nc);
  t->icell = (uint64_t *)malloc(sizeof(uint64_t) * nc);
  nodes2cells(t, root, 0, 1, X, R, 0, 0);
  free(root);
  return t;
}

void orc_tree_destroy(orc_tree *t) {
  if (!t) return;
  free(t->perm);
  free(t->iparent);
  free(t->ichild);
  free(t->nchild);
  free(t->ibody);
  free(t->nbody);
  free(t->X);
  free(t->R);
  free(t->icell);
  free(t);
}

int orc_tree_ncell(const orc_tree *t) { return t->ncell; }
int orc_tree_depth(const orc_tree *t) { return t->depth; }

void orc_tree_get(const orc_tree *t, int *perm, int *iparent, int *ichild, int *nchild,
                  int *ibody, int *nbody, double *X, double *R, uint64_t *icell) {
  int nc = t->ncell;
  if (perm) memcpy(perm, t->perm, sizeof(int) * t->n);
  if (iparent) memcpy(iparent, t->iparent, sizeof(int) * nc);
  if (ichild) memcpy(ichild, t->ichild, sizeof(int) * nc);
  if (nchild) memcpy(nchild, t->nchild, sizeof(int) * nc);
  if (ibody) memcpy(ibody, t->ibody, sizeof(int) * nc);
  if (nbody) memcpy(nbody, t->nbody, sizeof(int) * nc);
  if (X) memcpy(X, t->X, sizeof(double) * 3 * nc);
  if (R) memcpy(R, t->R, sizeof(double) * nc);
  if (icell) memcpy(icell, t->icell, sizeof(uint64_t) * nc);
}

/* ---------------- traversal (traversal.h) ---------------- */
struct orc_lists {
  int64_t n, cap, count[2];
  int *ci, *cj;
  unsigned char *type;
};

/* t: the target tree (Ci), tj: the source tree (Cj); the same tree for
 * traverse(cells, cells), two trees for traverse(icells, jcells) */
typedef struct {
  const orc_tree *t, *tj;
  double theta;
  int nspawn;
  double xper[3];
  orc_lists *out;
  /* when eval is set the pairs are evaluated immediately instead */
  const orc_kernel *k;
  const double *x, *q;    /* target bodies (and, one tree, the sources) */
  const double *xj, *qj;  /* source bodies */
  double *trg;
  cplx *M, *L;            /* M of the source tree, L of the target tree */
} trav_ctx;

static void emit(trav_ctx *c, int ci, int cj, int type) {
  if (c->k) {
    const orc_tree *t = c->t, *tj = c->tj;
    if (type == 0) {
      orc_m2l(c->k, t->X + 3 * ci, tj->X + 3 * cj, c->xper, (const double *)(c->M + (size_t)cj * c->k->NTERM),
              (double *)(c->L + (size_t)ci * c->k->NTERM));
    } else {
      int bi = t->ibody[ci], bj = tj->ibody[cj];
      orc_p2p(c->k, t->nbody[ci], c->x + 3 * (size_t)bi, c->trg + 4 * (size_t)bi, tj->nbody[cj],
              c->xj + 3 * (size_t)bj, c->qj + bj, c->xper);
    }
  }
  orc_lists *l = c->out;
  if (!l) return;
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 1024;
    l->ci = (int *)realloc(l->ci, sizeof(int) * l->cap);
    l->cj = (int *)realloc(l->cj, sizeof(int) * l->cap);
    l->type = (unsigned char *)realloc(l->type, l->cap);
  }
  l->ci[l->n] = ci;
  l->cj[l->n] = cj;
  l->type[l->n] = (unsigned char)type;
  l->n++;
  l->count[type]++;
}

static void dual_tree(trav_ctx *c, int Ci, int Cj);

/* TraverseRange::operator(), run serially in task order: traversal.h:196-229
 * (ranges of two different trees never begin at the same cell) */
static void traverse_range(trav_ctx *c, int ib, int ie, int jb, int je) {
  if (ie - ib == 1 || je - jb == 1) {
    if (ib == jb && c->t == c->tj) {
      dual_tree(c, ib, jb);
    } else {
      for (int i = ib; i < ie; i++)
        for (int j = jb; j < je; j++) dual_tree(c, i, j);
    }
  } else {
    int im = ib + (ie - ib) / 2, jm = jb + (je - jb) / 2;
    traverse_range(c, ib, im, jb, jm);
    traverse_range(c, im, ie, jm, je);
    traverse_range(c, ib, im, jm, je);
    traverse_range(c, im, ie, jb, jm);
  }
}

/* splitCell: traversal.h:90-114 */
static void split_cell(trav_ctx *c, int Ci, int Cj) {
  const orc_tree *t = c->t, *tj = c->tj;
  if (tj->nchild[Cj] == 0) {
    for (int ci = t->ichild[Ci]; ci < t->ichild[Ci] + t->nchild[Ci]; ci++) dual_tree(c, ci, Cj);
  } else if (t->nchild[Ci] == 0) {
    for (int cj = tj->ichild[Cj]; cj < tj->ichild[Cj] + tj->nchild[Cj]; cj++) dual_tree(c, Ci, cj);
  } else if (t->nbody[Ci] + tj->nbody[Cj] >= c->nspawn || (t == tj && Ci == Cj)) {
    traverse_range(c, t->ichild[Ci], t->ichild[Ci] + t->nchild[Ci], tj->ichild[Cj],
                   tj->ichild[Cj] + tj->nchild[Cj]);
  } else if (t->R[Ci] >= tj->R[Cj]) {
    for (int ci = t->ichild[Ci]; ci < t->ichild[Ci] + t->nchild[Ci]; ci++) dual_tree(c, ci, Cj);
  } else {
    for (int cj = tj->ichild[Cj]; cj < tj->ichild[Cj] + tj->nchild[Cj]; cj++) dual_tree(c, Ci, cj);
  }
}

/* dualTreeTraversal: traversal.h:117-177 (EXAFMM_NO_P2P off) */
static void dual_tree(trav_ctx *c, int Ci, int Cj) {
  const orc_tree *t = c->t, *tj = c->tj;
  double dX[3];
  for (int d = 0; d < 3; d++) dX[d] = t->X[3 * Ci + d] - tj->X[3 * Cj + d] - c->xper[d];
  double RT2 = norm3(dX) * c->theta * c->theta;
  double Rs = t->R[Ci] + tj->R[Cj];
  if (RT2 > Rs * Rs * (1 - 1e-3)) {
    emit(c, Ci, Cj, 0);
  } else if (t->nchild[Ci] == 0 && tj->nchild[Cj] == 0) {
    if (tj->nbody[Cj] == 0)
      emit(c, Ci, Cj, 0);
    else
      emit(c, Ci, Cj, 1);
  } else {
    split_cell(c, Ci, Cj);
  }
}

orc_lists *orc_traverse(const orc_tree *t, double theta, int nspawn, const double *xper) {
  orc_lists *l = (orc_lists *)calloc(1, sizeof(orc_lists));
  if (t->ncell == 0) return l;
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = c.tj = t;
  c.theta = theta;
  c.nspawn = nspawn;
  if (xper) memcpy(c.xper, xper, sizeof c.xper);
  c.out = l;
  dual_tree(&c, 0, 0);
  return l;
}

/* Traversal::traverse(icells, jcells) with icells != jcells
 * (include/traversal.h:326-333, examples/fmm.cxx:69-72): the dual traversal
 * from the pair of roots of two trees; ci indexes ti, cj indexes tj */
orc_lists *orc_traverse_pair(const orc_tree *ti, const orc_tree *tj, double theta, int nspawn,
                             const double *xper) {
  orc_lists *l = (orc_lists *)calloc(1, sizeof(orc_lists));
  if (ti->ncell == 0 || tj->ncell == 0) return l;
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = ti;
  c.tj = tj;
  c.theta = theta;
  c.nspawn = nspawn;
  if (xper) memcpy(c.xper, xper, sizeof c.xper);
  c.out = l;
  dual_tree(&c, 0, 0);
  return l;
}

void orc_lists_destroy(orc_lists *l) {
  if (!l) return;
  free(l->ci);
  free(l->cj);
  free(l->type);
  free(l);
}

int64_t orc_lists_count(const orc_lists *l, int type) {
  return type < 0 ? l->n : l->count[type];
}

void orc_lists_get(const orc_lists *l, int *ci, int *cj, int *type) {
  for (int64_t i = 0; i < l->n; i++) {
    if (ci) ci[i] = l->ci[i];
    if (cj) cj[i] = l->cj[i];
    if (type) type[i] = l->type[i];
  }
}

/* ---------------- up/down pass (up_down_pass.h) ---------------- */
typedef struct {
  const orc_kernel *k;
  const orc_tree *t;
  const double *x, *q;
  double *trg;
  cplx *M, *L;
} pass_ctx;

/* postOrderTraversal: up_down_pass.h:14-22 */
static void post_order(pass_ctx *c, int C) {
  const orc_tree *t = c->t;
  int NT = c->k->NTERM;
  for (int cc = t->ichild[C]; cc < t->ichild[C] + t->nchild[C]; cc++) post_order(c, cc);
  if (t->nchild[C] == 0) {
    int b = t->ibody[C];
    orc_p2m(c->k, t->X + 3 * C, t->nbody[C], c->x + 3 * (size_t)b, c->q + b,
            (double *)(c->M + (size_t)C * NT));
  } else {
    for (int cj = t->ichild[C]; cj < t->ichild[C] + t->nchild[C]; cj++)
      orc_m2m(c->k, t->X + 3 * C, t->X + 3 * cj, (const double *)(c->M + (size_t)cj * NT),
              (double *)(c->M + (size_t)C * NT));
  }
}

/* preOrderTraversal: up_down_pass.h:25-42 */
static void pre_order(pass_ctx *c, int C) {
  const orc_tree *t = c->t;
  int NT = c->k->NTERM;
  int Pc = t->iparent[C];
  orc_l2l(c->k, t->X + 3 * C, t->X + 3 * Pc, (const double *)(c->L + (size_t)Pc * NT),
          (double *)(c->L + (size_t)C * NT));
  if (t->nchild[C] == 0) {
    int b = t->ibody[C];
    orc_l2p(c->k, t->X + 3 * C, (const double *)(c->L + (size_t)C * NT), t->nbody[C],
            c->x + 3 * (size_t)b, c->trg + 4 * (size_t)b);
  }
  for (int cc = t->ichild[C]; cc < t->ichild[C] + t->nchild[C]; cc++) pre_order(c, cc);
}

void orc_upward(const orc_kernel *k, const orc_tree *t, const double *x, const double *q,
                double *M) {
  if (t->ncell == 0) return;
  memset(M, 0, sizeof(cplx) * (size_t)t->ncell * k->NTERM);
  pass_ctx c = {k, t, x, q, NULL, (cplx *)M, NULL};
  post_order(&c, 0);
}

/* downwardPass: up_down_pass.h:63-75 */
void orc_downward(const orc_kernel *k, const orc_tree *t, const double *x, double *L,
                  double *trg) {
  if (t->ncell == 0) return;
  pass_ctx c = {k, t, x, NULL, trg, NULL, (cplx *)L};
  if (t->nchild[0] == 0)
    orc_l2p(k, t->X, L, t->nbody[0], x + 3 * (size_t)t->ibody[0], trg + 4 * (size_t)t->ibody[0]);
  for (int cc = t->ichild[0]; cc < t->ichild[0] + t->nchild[0]; cc++) pre_order(&c, cc);
}

void orc_fmm(const orc_kernel *k, const orc_tree *t, const double *x, const double *q,
             double theta, int nspawn, double *trg, double *Mout, double *Lout) {
  if (t->ncell == 0) return;
  size_t sz = sizeof(cplx) * (size_t)t->ncell * k->NTERM;
  cplx *M = Mout ? (cplx *)Mout : (cplx *)malloc(sz);
  cplx *L = Lout ? (cplx *)Lout : (cplx *)malloc(sz);
  orc_upward(k, t, x, q, (double *)M);
  memset(L, 0, sz);
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = c.tj = t;
  c.theta = theta;
  c.nspawn = nspawn;
  c.k = k;
  c.x = c.xj = x;
  c.q = c.qj = q;
  c.trg = trg;
  c.M = M;
  c.L = L;
  dual_tree(&c, 0, 0);
  orc_downward(k, t, x, (double *)L, trg);
  if (!Mout) free(M);
  if (!Lout) free(L);
}

/* examples/fmm.cxx:59-77 with args.IneJ: upward pass of both trees, the
 * pair traversal (M2L from the source tree's multipoles into the target
 * tree's locals, P2P from source leaves to target leaves), downward pass of
 * the target tree.  x, q: target bodies; xj, qj: source bodies (tree order). */
void orc_fmm_pair(const orc_kernel *k, const orc_tree *ti, const double *x, const orc_tree *tj,
                  const double *xj, const double *qj, double theta, int nspawn, double *trg) {
  if (ti->ncell == 0 || tj->ncell == 0) return;
  const size_t szi = sizeof(cplx) * (size_t)ti->ncell * k->NTERM;
  const size_t szj = sizeof(cplx) * (size_t)tj->ncell * k->NTERM;
  cplx *M = (cplx *)malloc(szj), *L = (cplx *)calloc(1, szi);
  orc_upward(k, tj, xj, qj, (double *)M);
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = ti;
  c.tj = tj;
  c.theta = theta;
  c.nspawn = nspawn;
  c.k = k;
  c.x = x;
  c.xj = xj;
  c.qj = qj;
  c.trg = trg;
  c.M = M;
  c.L = L;
  dual_tree(&c, 0, 0);
  orc_downward(k, ti, x, (double *)L, trg);
  free(M);
  free(L);
}

/* Traversal::traversePeriodic  include/traversal.h:233-287 (target Ci0 = the
 * root, source Cj0 = the root of the same tree) */
static void traverse_periodic(const orc_kernel *k, const orc_tree *t, double theta, int images,
                              const double *cycle0, const cplx *M, cplx *L) {
  const int NT = k->NTERM;
  const int prange = (int)(.5 / theta + 1), neighbor = 2 * prange + 1;
  cplx *Mci = (cplx *)malloc(sizeof(cplx) * NT), *Mcopy = (cplx *)malloc(sizeof(cplx) * NT);
  memcpy(Mci, M, sizeof(cplx) * NT); /* *Ci = *Cj0 */
  const double *X = t->X;            /* Ci->X = Cj0->X = Ci0->X */
  double cycle[3] = {cycle0[0], cycle0[1], cycle0[2]};
  for (int level = 0; level < images - 1; level++) {
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          if (ix == 0 && iy == 0 && iz == 0) continue;
          for (int cx = -prange; cx <= prange; cx++)
            for (int cy = -prange; cy <= prange; cy++)
              for (int cz = -prange; cz <= prange; cz++) {
                double xper[3] = {(ix * neighbor + cx) * cycle[0], (iy * neighbor + cy) * cycle[1],
                                  (iz * neighbor + cz) * cycle[2]};
                orc_m2l(k, X, X, xper, (const double *)Mci, (double *)L);
              }
        }
    /* the image copies (children of Ci) take Ci's multipole before the M2M */
    memcpy(Mcopy, Mci, sizeof(cplx) * NT);
    for (int ix = -prange; ix <= prange; ix++)
      for (int iy = -prange; iy <= prange; iy++)
        for (int iz = -prange; iz <= prange; iz++) {
          if (ix == 0 && iy == 0 && iz == 0) continue;
          double Xc[3] = {X[0] + ix * cycle[0], X[1] + iy * cycle[1], X[2] + iz * cycle[2]};
          orc_m2m(k, X, Xc, (const double *)Mcopy, (double *)Mci);
        }
    for (int d = 0; d < 3; d++) cycle[d] *= neighbor;
  }
  free(Mci);
  free(Mcopy);
}

void orc_fmm_periodic(const orc_kernel *k, const orc_tree *t, const double *x, const double *q,
                      double theta, int nspawn, int images, const double *cycle, double *trg) {
  if (t->ncell == 0) return;
  size_t sz = sizeof(cplx) * (size_t)t->ncell * k->NTERM;
  cplx *M = (cplx *)malloc(sz), *L = (cplx *)malloc(sz);
  orc_upward(k, t, x, q, (double *)M);
  memset(L, 0, sz);
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = c.tj = t;
  c.theta = theta;
  c.nspawn = nspawn;
  c.k = k;
  c.x = c.xj = x;
  c.q = c.qj = q;
  c.trg = trg;
  c.M = M;
  c.L = L;
  if (images == 0) {
    dual_tree(&c, 0, 0);
  } else {
    for (int ix = -1; ix <= 1; ix++)
      for (int iy = -1; iy <= 1; iy++)
        for (int iz = -1; iz <= 1; iz++) {
          c.xper[0] = ix * cycle[0];
          c.xper[1] = iy * cycle[1];
          c.xper[2] = iz * cycle[2];
          dual_tree(&c, 0, 0);
        }
    traverse_periodic(k, t, theta, images, cycle, M, L);
  }
  orc_downward(k, t, x, (double *)L, trg);
  free(M);
  free(L);
}

void orc_eval_lists(const orc_kernel *k, const orc_tree *t, const double *x, const double *q,
                    int64_t npair, const int *ci, const int *cj, const int *type, double *trg,
                    double *L, const double *M) {
  trav_ctx c;
  memset(&c, 0, sizeof c);
  c.t = c.tj = t;
  c.k = k;
  c.x = c.xj = x;
  c.q = c.qj = q;
  c.trg = trg;
  c.M = (cplx *)M;
  c.L = (cplx *)L;
  for (int64_t i = 0; i < npair; i++) emit(&c, ci[i], cj[i], type[i]);
}

void orc_direct(const orc_kernel *k, int ni, const double *xi, double *trgi, int nj,
                const double *xj, const double *qj) {
  double xper[3] = {0, 0, 0};
  orc_p2p(k, ni, xi, trgi, nj, xj, qj, xper);
}

/* verify.h:42-92 and examples/fmm.cxx:99-104 */
void orc_rel_l2(int n, const double *ref, const double *trg, double *pot_rel, double *acc_rel) {
  double pd = 0, pn = 0, ad = 0, an = 0;
  for (int i = 0; i < n; i++) {
    const double *a = ref + 4 * i, *b = trg + 4 * i;
    pd += fabs((a[0] - b[0]) * (a[0] - b[0]));
    pn += fabs(a[0] * a[0]);
    ad += fabs((a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]) +
               (a[3] - b[3]) * (a[3] - b[3]));
    an += fabs(a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
  }
  *pot_rel = sqrt(pd / pn);
  *acc_rel = sqrt(ad / an);
}

/* ---------------- datasets (dataset.h) ---------------- */
void orc_data_cube(int n, int seed, double *x) {
  srand48(seed);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) x[3 * i + d] = drand48() * 2 * M_PI - M_PI;
}

void orc_data_plummer(int n, int seed, double *x) {
  srand48(seed);
  int i = 0;
  while (i < n) {
    double X1 = drand48(), X2 = drand48(), X3 = drand48();
    double R = 1.0 / sqrt((pow(X1, -2.0 / 3.0) - 1.0));
    if (R < 100.0) {
      double Z = (1.0 - 2.0 * X2) * R;
      double X = sqrt(R * R - Z * Z) * cos(2.0 * M_PI * X3);
      double Y = sqrt(R * R - Z * Z) * sin(2.0 * M_PI * X3);
      double scale = 3.0 * M_PI / 16.0;
      X *= scale;
      Y *= scale;
      Z *= scale;
      x[3 * i + 0] = X;
      x[3 * i + 1] = Y;
      x[3 * i + 2] = Z;
      i++;
    }
  }
}

void orc_data_source(int n, int seed, double *q) {
  srand48(seed);
  double average = 0;
  for (int i = 0; i < n; i++) {
    q[i] = drand48() - .5;
    average += q[i];
  }
  average /= n;
  for (int i = 0; i < n; i++) q[i] -= average;
}

int orc_sample_indices(int n, int ntarget, int *idx) {
  if (ntarget < n) {
    int stride = n / ntarget;
    for (int i = 0; i < ntarget; i++) idx[i] = i * stride;
    return ntarget;
  }
  for (int i = 0; i < n; i++) idx[i] = i;
  return n;
}

/* ---------------------------------------------------------------- periodic MD pieces */

/* vec.h:217-230: wrap one coordinate triple into [-cycle/2, cycle/2]; returns
 * the bit mask of the axes moved (unwrap, vec.h:231-235, reverses it). */
int orc_wrap(double *x, const double *cycle) {
  int iw = 0;
  for (int d = 0; d < 3; d++) {
    if (x[d] < -cycle[d] / 2) {
      x[d] += cycle[d];
      iw |= 1 << d;
    }
    if (x[d] > cycle[d] / 2) {
      x[d] -= cycle[d];
      iw |= 1 << d;
    }
  }
  return iw;
}

/* UpDownPass::getDipole(bodies, X0 = 0) and dipoleCorrection,
 * include/up_down_pass.h:77-99: trg -= 4 pi / (3 V) * (|D|^2 / (N q_i), D). */
void orc_dipole_correction(int n, const double *x, const double *q, const double *cycle, double *trg) {
  double D[3] = {0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) D[d] += x[3 * i + d] * q[i];
  const double coef = 4 * M_PI / (3 * cycle[0] * cycle[1] * cycle[2]);
  const double dd = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  for (int i = 0; i < n; i++) {
    trg[4 * i] -= coef * dd / n / q[i];
    for (int d = 0; d < 3; d++) trg[4 * i + 1 + d] -= coef * D[d];
  }
}

/* Ewald summation, include/ewald.h (tin-foil boundary):
 *   wave part  (initWaves :64-88, dft :26-40, factor :163-180, idft :43-60):
 *     k = (l, m, n) over the half space l > 0 | (l = 0, m > 0) | (l = m = 0, n > 0)
 *     with l^2 + m^2 + n^2 <= ksize^2; S(k) = sum_j q_j e^{i k.x_j 2pi/L};
 *     phi_i += 2/(sigma V) sum_k e^{-K^2/4 alpha^2}/K^2 Re(S(k) e^{-i K.x_i})
 *   real part  (P2P :90-110): every pair within the cutoff,
 *     phi_i += q_j erfc(alpha r)/r, grad_i -= dX q_j (2/sqrt(pi) e^{-a^2 r^2}/(a r)^2
 *     + erfc(a r)/(a r)^3) a^3
 *   self term  (:156-160): phi_i -= 2/sqrt(pi) q_i alpha.
 * The reference finds real-part pairs with a cell search that gives every
 * (target leaf, source leaf) pair the image that wraps their centres
 * (Neighbor, :113-137); here every body pair takes its own minimum image,
 * which is the sum that search is meant to produce (they differ only by
 * pairs the search places across the box boundary).  trg += the result. */
void orc_ewald(int n, const double *x, const double *q, int ksize, double alpha, double sigma, double cutoff,
               const double *cycle, double *trg) {
  double scale[3];
  for (int d = 0; d < 3; d++) scale[d] = 2 * M_PI / cycle[d];
  const double coef = 2 / sigma / cycle[0] / cycle[1] / cycle[2];
  const double coef2 = 1 / (4 * alpha * alpha);
  const int kmax = ksize, kmaxsq = ksize * ksize;
  for (int l = 0; l <= kmax; l++) {
    const int mmin = l == 0 ? 0 : -kmax;
    for (int m = mmin; m <= kmax; m++) {
      const int nmin = (l == 0 && m == 0) ? 1 : -kmax;
      for (int nn = nmin; nn <= kmax; nn++) {
        if (l * l + m * m + nn * nn > kmaxsq) continue;
        const double K[3] = {l, m, nn};
        double re = 0, im = 0;
        for (int j = 0; j < n; j++) {
          double th = 0;
          for (int d = 0; d < 3; d++) th += K[d] * x[3 * j + d] * scale[d];
          re += q[j] * cos(th);
          im += q[j] * sin(th);
        }
        double K2 = 0;
        for (int d = 0; d < 3; d++) K2 += (K[d] * scale[d]) * (K[d] * scale[d]);
        const double factor = coef * exp(-K2 * coef2) / K2;
        re *= factor;
        im *= factor;
        for (int i = 0; i < n; i++) {
          double th = 0;
          for (int d = 0; d < 3; d++) th += K[d] * x[3 * i + d] * scale[d];
          const double c = cos(th), s = sin(th);
          const double dtmp = re * s - im * c;
          trg[4 * i] += re * c + im * s;
          for (int d = 0; d < 3; d++) trg[4 * i + 1 + d] -= dtmp * K[d] * scale[d];
        }
      }
    }
  }
  const double c2 = cutoff * cutoff;
  for (int i = 0; i < n; i++) {
    double p = 0, g[3] = {0, 0, 0};
    for (int j = 0; j < n; j++) {
      double dX[3];
      for (int d = 0; d < 3; d++) dX[d] = x[3 * i + d] - x[3 * j + d];
      orc_wrap(dX, cycle);
      const double R2 = dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2];
      if (0 < R2 && R2 < c2) {
        const double R2s = R2 * alpha * alpha, Rs = sqrt(R2s);
        const double invRs = 1 / Rs, invR2s = invRs * invRs, invR3s = invR2s * invRs;
        double dtmp = q[j] * (M_2_SQRTPI * exp(-R2s) * invR2s + erfc(Rs) * invR3s);
        dtmp *= alpha * alpha * alpha;
        p += q[j] * erfc(Rs) * invRs * alpha;
        for (int d = 0; d < 3; d++) g[d] -= dX[d] * dtmp;
      }
    }
    trg[4 * i] += p - M_2_SQRTPI * q[i] * alpha;
    for (int d = 0; d < 3; d++) trg[4 * i + 1 + d] += g[d];
  }
}
