/* CPU baseline and cross-check (test infrastructure, not product code):
 * one red-black overrelaxation half sweep on the reference's host layout
 * (gauge field gf[4 s + mu], transformation trafo[s], full 3x3 su3 of
 * `complex`, global.h:26-33), built from the reference's own host SU(3)
 * and complex routines (su3manip.c: su3zero, su3mult, su3dagger, su3add,
 * su3sub, su3copy; complex.c: cmult, cadd, csub, cconj, crealmult,
 * initcomplex).  The site update restates dev_overrelax_step
 * (overrelax.cu:335-446): star K = sum_mu U_mu(x) g^+(x+mu) +
 * U_mu^+(x-mu) g^+(x-mu); h = g(x) relaxed in the three Cabibbo-Marinari
 * subgroups (overrelax.cu:188-269) and normalised; g(x) <- normalize(R_w)
 * g(x) with R_w = 1 + C(w,1) R + C(w,2) R^2, R = h g^+ - 1
 * (overrelax.cu:149-177); g(x) normalised (dev_su3.cu:1569-1634, with the
 * stored row 2 rebuilt as conj(row0 x row1)).  Built into oracle/_ref only,
 * where the reference checkout is present. */
#include <math.h>
#include <stddef.h>

/* the reference's headers, in the order its own host files include them */
#include "cudaglobal.h"
#include "global.h"
#include "complex.h"
#include "su3manip.h"

/* dev_su3.cu:1569-1634: rows 0, 1 Gram-Schmidt, row 2 = conj(row0 x row1) */
static void ref_normalize(su3 *u) {
  double len = 0.0;
  int i;
  complex proj = initcomplex(0.0, 0.0);
  for (i = 0; i < 3; i++) len += (*u)[0][i].re * (*u)[0][i].re + (*u)[0][i].im * (*u)[0][i].im;
  len = 1.0 / sqrt(len);
  for (i = 0; i < 3; i++) (*u)[0][i] = crealmult((*u)[0][i], len);
  for (i = 0; i < 3; i++) proj = cadd(proj, cmult((*u)[1][i], cconj((*u)[0][i])));
  for (i = 0; i < 3; i++) (*u)[1][i] = csub((*u)[1][i], cmult(proj, (*u)[0][i]));
  len = 0.0;
  for (i = 0; i < 3; i++) len += (*u)[1][i].re * (*u)[1][i].re + (*u)[1][i].im * (*u)[1][i].im;
  len = 1.0 / sqrt(len);
  for (i = 0; i < 3; i++) (*u)[1][i] = crealmult((*u)[1][i], len);
  (*u)[2][0] = cconj(csub(cmult((*u)[0][1], (*u)[1][2]), cmult((*u)[0][2], (*u)[1][1])));
  (*u)[2][1] = cconj(csub(cmult((*u)[0][2], (*u)[1][0]), cmult((*u)[0][0], (*u)[1][2])));
  (*u)[2][2] = cconj(csub(cmult((*u)[0][0], (*u)[1][1]), cmult((*u)[0][1], (*u)[1][0])));
}

/* overrelax.cu:188-269: for the subgroups (a,b), X = h K; w from X; h's
 * rows a, b <- [[w0 + i w3, w2 + i w1], [-w2 + i w1, w0 - i w3]] rows */
static void ref_cm_relax(su3 *h, su3 *K) {
  su3 X;
  int a, b, c;
  for (a = 0; a < 2; a++)
    for (b = a + 1; b < 3; b++) {
      double w0, w1, w2, w3, n;
      su3mult(&X, h, K);
      w0 = X[a][a].re + X[b][b].re;
      w3 = -X[a][a].im + X[b][b].im;
      w1 = -X[a][b].im - X[b][a].im;
      w2 = -X[a][b].re + X[b][a].re;
      n = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2 + w3 * w3);
      w0 *= n; w1 *= n; w2 *= n; w3 *= n;
      for (c = 0; c < 3; c++) {
        complex ra = cadd(cmult(initcomplex(w0, w3), (*h)[a][c]), cmult(initcomplex(w2, w1), (*h)[b][c]));
        complex rb = cadd(cmult(initcomplex(-w2, w1), (*h)[a][c]), cmult(initcomplex(w0, -w3), (*h)[b][c]));
        (*h)[a][c] = ra;
        (*h)[b][c] = rb;
      }
    }
}

/* overrelax.cu:149-177 with its three terms: g <- normalize(1 + c1 R +
 * c2 R^2) g, R = h g^+ - 1, c_i = (1 + w - i) / i applied cumulatively */
static void ref_overrelax(su3 *g, su3 *h, double omega) {
  su3 one, gd, R, a, help, Rw;
  int i, p, q;
  su3zero(&one);
  for (i = 0; i < 3; i++) one[i][i] = initcomplex(1.0, 0.0);
  su3dagger(&gd, g);
  su3mult(&R, h, &gd);
  su3sub(&R, &R, &one);
  su3copy(&Rw, &one);
  su3copy(&a, &one);
  for (i = 1; i < 3; i++) {
    su3mult(&help, &a, &R);
    for (p = 0; p < 3; p++)
      for (q = 0; q < 3; q++) a[p][q] = crealmult(help[p][q], (1.0 + omega - i) / i);
    su3add(&Rw, &Rw, &a);
  }
  ref_normalize(&Rw);
  su3mult(&help, &Rw, g);
  su3copy(g, &help);
}

static void ref_site(su3 *gf, su3 *trafo, const int *nn, int pos, int ndir, double omega) {
  su3 K, A, B, P, h;
  int mu;
  su3zero(&K);
  for (mu = 0; mu < ndir; mu++) {
    int hop = nn[8 * pos + mu];
    su3dagger(&B, &trafo[hop]);
    su3mult(&P, &gf[4 * pos + mu], &B);
    su3add(&K, &K, &P);
    hop = nn[8 * pos + 4 + mu];
    su3dagger(&A, &gf[4 * hop + mu]);
    su3dagger(&B, &trafo[hop]);
    su3mult(&P, &A, &B);
    su3add(&K, &K, &P);
  }
  su3copy(&h, &trafo[pos]);
  ref_cm_relax(&h, &K);
  ref_normalize(&h);
  ref_overrelax(&trafo[pos], &h, omega);
  ref_normalize(&trafo[pos]);
}

/* the sites[0 .. nsites) of one parity; nn[8 s + mu] forward, nn[8 s + 4 +
 * mu] backward neighbours (cudagaugefix.cu:72-157 order) */
void ref_or_halfsweep(su3 *gf, su3 *trafo, const int *nn, const int *sites, int nsites, int ndir,
                      double omega) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < nsites; i++) ref_site(gf, trafo, nn, sites[i], ndir, omega);
}
