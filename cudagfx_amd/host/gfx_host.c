/*
 * gfx_host.c -- host services around the device path: the Lehmer RNG that
 * seeds the initial gauge transformation, the random/unit initialisation of
 * g and U, and the host observables the driver prints next to the device
 * values.
 *
 * Neighbours are computed from the coordinates (periodic, mu = x,y,z,t)
 * rather than from a stored table.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/cudagfx_host.h"

/* ------------------------------------------------------------------ Lehmer RNG
 * Park & Miller multiplicative congruential generator, modulus 2^31-1,
 * multiplier 48271, 256 streams spaced by the jump multiplier 22925
 * (rngs.c:39-43); Schrage's factorisation avoids 64-bit overflow. */
enum { LEH_STREAMS = 256 };
static const long LEH_M = 2147483647L;
static const long LEH_A = 48271L;
static const long LEH_JUMP = 22925L;
static const long LEH_DEFAULT = 123456789L;
static long leh_state[LEH_STREAMS] = {123456789L};
static int leh_cur = 0;
static int leh_planted = 0;

static long leh_mul(long a, long x) {
  const long q = LEH_M / a, r = LEH_M % a;
  const long t = a * (x % q) - r * (x / q);
  return t > 0 ? t : t + LEH_M;
}

double Random(void) {
  leh_state[leh_cur] = leh_mul(LEH_A, leh_state[leh_cur]);
  return (double)leh_state[leh_cur] / LEH_M;
}

void PutSeed(long x) {
  if (x > 0) x = x % LEH_M;
  if (x < 0) x = ((unsigned long)time(NULL)) % LEH_M;
  while (x == 0) {
    printf("\nEnter a positive integer seed (9 digits or less) >> ");
    if (scanf("%ld", &x) != 1) x = LEH_DEFAULT;
    if (!(0 < x && x < LEH_M)) {
      printf("\nInput out of range ... try again\n");
      x = 0;
    }
  }
  leh_state[leh_cur] = x;
}

void GetSeed(long *x) { *x = leh_state[leh_cur]; }

void PlantSeeds(long x) {
  const int keep = leh_cur;
  leh_planted = 1;
  leh_cur = 0;
  PutSeed(x);
  leh_cur = keep;
  for (int j = 1; j < LEH_STREAMS; j++) leh_state[j] = leh_mul(LEH_JUMP, leh_state[j - 1]);
}

void SelectStream(int index) {
  leh_cur = (int)(((unsigned int)index) % LEH_STREAMS);
  if (!leh_planted && leh_cur != 0) PlantSeeds(LEH_DEFAULT);
}

/* ------------------------------------------------------------------ SU(3) helpers */
typedef struct {
  double re, im;
} cx_t;

static inline cx_t cxm(double re, double im) {
  cx_t c = {re, im};
  return c;
}
static inline cx_t cx_mul(cx_t a, cx_t b) {
  return cxm(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx_t cx_add(cx_t a, cx_t b) { return cxm(a.re + b.re, a.im + b.im); }
static inline cx_t cx_sub(cx_t a, cx_t b) { return cxm(a.re - b.re, a.im - b.im); }
static inline cx_t cx_conj(cx_t a) { return cxm(a.re, -a.im); }

typedef cx_t mat3[3][3];

static void mat_mul(mat3 r, mat3 a, mat3 b) {
  mat3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cx_t s = cxm(0.0, 0.0);
      for (int k = 0; k < 3; k++) s = cx_add(cx_mul(a[i][k], b[k][j]), s);
      t[i][j] = s;
    }
  memcpy(r, t, sizeof(mat3));
}
static void mat_dag(mat3 r, mat3 a) {
  mat3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = cx_conj(a[j][i]);
  memcpy(r, t, sizeof(mat3));
}

/* a random SU(3) matrix from ten Lehmer draws: two rows filled component by
 * component with the remaining norm, Gram-Schmidt on row 1, row 2 the
 * conjugated cross product (su3manip.c:539-641) */
static void random_su3_matrix(mat3 m) {
  double norm, h;
  for (int r = 0; r < 2; r++) {
    double v[6];
    norm = 0.0;
    h = Random();
    v[0] = h;
    norm += h * h;
    for (int k = 1; k < 5; k++) {
      h = (1.0 - norm) * Random();
      v[k] = h;
      norm += h * h;
    }
    v[5] = sqrt(1.0 - norm);
    for (int c = 0; c < 3; c++) m[r][c] = cxm(v[2 * c], v[2 * c + 1]);
  }
  cx_t proj = cxm(0.0, 0.0);
  for (int j = 0; j < 3; j++) proj = cx_add(proj, cx_mul(m[1][j], cx_conj(m[0][j])));
  for (int j = 0; j < 3; j++) m[1][j] = cx_sub(m[1][j], cx_mul(proj, m[0][j]));
  double len = m[1][0].re * m[1][0].re + m[1][0].im * m[1][0].im + m[1][1].re * m[1][1].re +
               m[1][1].im * m[1][1].im + m[1][2].re * m[1][2].re + m[1][2].im * m[1][2].im;
  len = 1.0 / sqrt(len);
  for (int j = 0; j < 3; j++) {
    m[1][j].re = m[1][j].re * len;
    m[1][j].im = m[1][j].im * len;
  }
  cx_t a;
  a = cx_sub(cx_mul(m[0][1], m[1][2]), cx_mul(m[0][2], m[1][1]));
  m[2][0] = cxm(a.re, -a.im);
  a = cx_sub(cx_mul(m[0][2], m[1][0]), cx_mul(m[0][0], m[1][2]));
  m[2][1] = cxm(a.re, -a.im);
  a = cx_sub(cx_mul(m[0][0], m[1][1]), cx_mul(m[0][1], m[1][0]));
  m[2][2] = cxm(a.re, -a.im);
}

static void unit_matrix(mat3 m) {
  memset(m, 0, sizeof(mat3));
  m[0][0].re = 1.0;
  m[1][1].re = 1.0;
  m[2][2].re = 1.0;
}

void unit_init_trafo(gfx_su3 *g) {
  for (long i = 0; i < VOLUME; i++) unit_matrix(*(mat3 *)&g[i]);
}
void unit_init_gauge(gfx_su3 *u) {
  for (long i = 0; i < 4L * VOLUME; i++) unit_matrix(*(mat3 *)&u[i]);
}
void random_init_trafo(gfx_su3 *g) {
  printf("Setting %d random gauge trafo fields\n", VOLUME);
  for (long i = 0; i < VOLUME; i++) random_su3_matrix(*(mat3 *)&g[i]);
}
void random_init_gauge(gfx_su3 *u) {
  printf("Setting %d random gauge fields\n", VOLUME);
  for (long i = 0; i < 4L * VOLUME; i++) random_su3_matrix(*(mat3 *)&u[i]);
}

/* ------------------------------------------------------------------ geometry */
static long nbr(long s, int mu, int dir) {
  const int d[4] = {LX, LY, LZ, T};
  long x[4], r = s;
  for (int k = 0; k < 4; k++) {
    x[k] = r % d[k];
    r /= d[k];
  }
  x[mu] = (x[mu] + dir + d[mu]) % d[mu];
  return x[0] + (long)LX * (x[1] + (long)LY * (x[2] + (long)LZ * x[3]));
}

/* ------------------------------------------------------------------ observables */
void g_trafo(gfx_su3 *u, gfx_su3 *g) {
  for (long s = 0; s < VOLUME; s++)
    for (int mu = 0; mu < 4; mu++) {
      mat3 h1, h2, *U = (mat3 *)&u[4 * s + mu];
      mat_mul(h1, *(mat3 *)&g[s], *U);
      mat_dag(h2, *(mat3 *)&g[nbr(s, mu, +1)]);
      mat_mul(*U, h1, h2);
    }
}

double gauge_functional(gfx_su3 *gf) {
  double h = 0.0;
  for (long k = 0; k < 4L * VOLUME; k++) {
    mat3 *m = (mat3 *)&gf[k];
    cx_t t = cx_add(cx_add(cx_add(cxm(0.0, 0.0), (*m)[0][0]), (*m)[1][1]), (*m)[2][2]);
    h += t.re;
  }
  return h / (4.0 * (double)VOLUME);
}

/* A = (-i/2)(U - U^+), traceless */
static void vecpot_host(mat3 A, mat3 U) {
  mat3 d;
  mat_dag(d, U);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] = cx_mul(cxm(0.0, -0.5), cx_sub(U[i][j], d[i][j]));
  cx_t tr = cx_add(cx_add(cx_add(cxm(0.0, 0.0), A[0][0]), A[1][1]), A[2][2]);
  tr.re = tr.re / 3.0;
  tr.im = tr.im / 3.0;
  for (int k = 0; k < 3; k++) A[k][k] = cx_sub(A[k][k], tr);
}

double dAdA(gfx_su3 *gf) {
  double sum = 0.0, mx = 0.0;
  for (long s = 0; s < VOLUME; s++) {
    mat3 dA, M1, M2, h1, h4;
    memset(dA, 0, sizeof(mat3));
    for (int mu = 0; mu < 4; mu++) {
      vecpot_host(M1, *(mat3 *)&gf[4 * nbr(s, mu, -1) + mu]);
      vecpot_host(M2, *(mat3 *)&gf[4 * s + mu]);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dA[i][j] = cx_add(cx_sub(M1[i][j], M2[i][j]), dA[i][j]);
    }
    mat_dag(h1, dA);
    mat_mul(h4, dA, h1);
    const double t = cx_add(cx_add(cx_add(cxm(0.0, 0.0), h4[0][0]), h4[1][1]), h4[2][2]).re;
    sum += t;
    if (t > mx) mx = t;
  }
  maxDADA = mx;
  return sum / (double)VOLUME;
}

double mean_plaq(gfx_su3 *gf) {
  double mp = 0.0;
  for (long x0 = 0; x0 < VOLUME; x0++)
    for (int nu = 0; nu < 3; nu++)
      for (int mu = nu + 1; mu < 4; mu++) {
        const long x1 = nbr(x0, mu, +1), x2 = nbr(x0, nu, +1);
        mat3 M3, M4, s1, s2;
        mat_dag(M3, *(mat3 *)&gf[4 * x2 + mu]);
        mat_dag(M4, *(mat3 *)&gf[4 * x0 + nu]);
        mat_mul(s1, M3, M4);
        mat_mul(s2, *(mat3 *)&gf[4 * x1 + nu], s1);
        mat_mul(s1, *(mat3 *)&gf[4 * x0 + mu], s2);
        mp += cx_add(cx_add(cx_add(cxm(0.0, 0.0), s1[0][0]), s1[1][1]), s1[2][2]).re / 3.0;
      }
  return mp * (1.0 / (6.0 * (double)VOLUME));
}
