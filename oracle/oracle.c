/*
 * oracle.c -- CPU restatement of the cudagfx gauge-fixing path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with
 * -ffp-contract=off so that every product and sum is rounded exactly as
 * written, in the order the reference writes it.  Citations are
 * reference-relative file:line.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cpx;
typedef cpx m3[3][3];

/* ---- complex scalar ops: dev_su3.cu:17-83 / complex.c ---- */
static inline cpx c_mk(double re, double im) { cpx c; c.re = re; c.im = im; return c; }
static inline cpx c_conj(cpx a) { return c_mk(a.re, -1.0 * a.im); }
static inline cpx c_mul(cpx a, cpx b) {
  return c_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cpx c_add(cpx a, cpx b) { return c_mk(a.re + b.re, a.im + b.im); }
static inline cpx c_sub(cpx a, cpx b) { return c_mk(a.re - b.re, a.im - b.im); }

/* ---- SU(3) helpers: dev_su3.cu:1321-1565 ---- */
static void m_zero(m3 m) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = c_mk(0.0, 0.0);
}
static void m_unit(m3 m) {
  m_zero(m);
  m[0][0].re = 1.0; m[1][1].re = 1.0; m[2][2].re = 1.0;
}
static void m_copy(m3 to, m3 from) { memcpy(to, from, sizeof(m3)); }
/* u = v * w (dev_su3.cu:1418-1433): help2 = cmult + help2, k = 0,1,2 */
static void m_mul(m3 u, m3 v, m3 w) {
  m3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = c_mk(0.0, 0.0);
      for (int k = 0; k < 3; k++) s = c_add(c_mul(v[i][k], w[k][j]), s);
      r[i][j] = s;
    }
  m_copy(u, r);
}
/* u += v * w (dev_su3.cu:1440-1456) */
static void m_addmul(m3 u, m3 v, m3 w) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = c_mk(0.0, 0.0);
      for (int k = 0; k < 3; k++) s = c_add(c_mul(v[i][k], w[k][j]), s);
      u[i][j].re += s.re;
      u[i][j].im += s.im;
    }
}
/* u = v * w^+ (dev_su3.cu:1461-1475) */
static void m_mul_d(m3 u, m3 v, m3 w) {
  m3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = c_mk(0.0, 0.0);
      for (int k = 0; k < 3; k++) s = c_add(c_mul(v[i][k], c_conj(w[j][k])), s);
      r[i][j] = s;
    }
  m_copy(u, r);
}
static void m_dagger(m3 out, m3 in) {
  m3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r[i][j] = c_conj(in[j][i]);
  m_copy(out, r);
}
static double m_retrace(m3 m) {   /* dev_su3.cu:1359-1367 */
  double e = 0.0;
  for (int i = 0; i < 3; i++) e = e + m[i][i].re;
  return e;
}
static cpx m_trace(m3 m) {        /* dev_su3.cu:1347-1355 */
  cpx e = c_mk(0.0, 0.0);
  for (int i = 0; i < 3; i++) e = c_add(e, m[i][i]);
  return e;
}

/* third row from the conjugated cross product of rows 0 and 1
 * (dev_su3.cu:1617-1632, also 800-817) */
static void m_row2(m3 u) {
  cpx h1, h2;
  h1 = c_mul(u[0][1], u[1][2]); h2 = c_mul(u[0][2], u[1][1]); u[2][0] = c_conj(c_sub(h1, h2));
  h1 = c_mul(u[0][2], u[1][0]); h2 = c_mul(u[0][0], u[1][2]); u[2][1] = c_conj(c_sub(h1, h2));
  h1 = c_mul(u[0][0], u[1][1]); h2 = c_mul(u[0][1], u[1][0]); u[2][2] = c_conj(c_sub(h1, h2));
}

/* dev_su3.cu:737-820 (dev_reconstructtrafo_2vtexref) */
static void m_load(m3 u, const double *r) {
  for (int j = 0; j < 3; j++) {
    u[0][j] = c_mk(r[2 * j], r[2 * j + 1]);
    u[1][j] = c_mk(r[6 + 2 * j], r[6 + 2 * j + 1]);
  }
  m_row2(u);
}
/* dev_su3.cu:826-929 (dev_reconstructtrafo_2vtexref_dagger) */
static void m_load_dagger(m3 u, const double *r) {
  cpx h1, h2;
  for (int i = 0; i < 3; i++) {
    u[i][0] = c_mk(r[2 * i], -r[2 * i + 1]);
    u[i][1] = c_mk(r[6 + 2 * i], -r[6 + 2 * i + 1]);
  }
  h1 = c_conj(c_mul(u[1][0], u[2][1])); h2 = c_conj(c_mul(u[2][0], u[1][1])); u[0][2] = c_sub(h1, h2);
  h1 = c_conj(c_mul(u[2][0], u[0][1])); h2 = c_conj(c_mul(u[0][0], u[2][1])); u[1][2] = c_sub(h1, h2);
  h1 = c_conj(c_mul(u[0][0], u[1][1])); h2 = c_conj(c_mul(u[1][0], u[0][1])); u[2][2] = c_sub(h1, h2);
}
/* dev_su3.cu:118-135 (dev_storetrafo_2v) */
static void m_store(double *r, m3 u) {
  for (int j = 0; j < 3; j++) {
    r[2 * j] = u[0][j].re; r[2 * j + 1] = u[0][j].im;
    r[6 + 2 * j] = u[1][j].re; r[6 + 2 * j + 1] = u[1][j].im;
  }
}

/* dev_su3.cu:1569-1634 (dev_su3_normalize); rsqrt restated as 1/sqrt */
static void m_normalize(m3 u) {
  double len;
  cpx proj;
  len = (u[0][0].re * u[0][0].re + u[0][0].im * u[0][0].im) +
        (u[0][1].re * u[0][1].re + u[0][1].im * u[0][1].im) +
        (u[0][2].re * u[0][2].re + u[0][2].im * u[0][2].im);
  len = 1.0 / sqrt(len);
  for (int i = 0; i < 3; i++) { u[0][i].re = u[0][i].re * len; u[0][i].im = u[0][i].im * len; }
  proj = c_mk(0.0, 0.0);
  for (int i = 0; i < 3; i++) proj = c_add(proj, c_mul(u[1][i], c_conj(u[0][i])));
  for (int i = 0; i < 3; i++) u[1][i] = c_sub(u[1][i], c_mul(proj, u[0][i]));
  len = (u[1][0].re * u[1][0].re + u[1][0].im * u[1][0].im) +
        (u[1][1].re * u[1][1].re + u[1][1].im * u[1][1].im) +
        (u[1][2].re * u[1][2].re + u[1][2].im * u[1][2].im);
  len = 1.0 / sqrt(len);
  for (int i = 0; i < 3; i++) { u[1][i].re = u[1][i].re * len; u[1][i].im = u[1][i].im * len; }
  m_row2(u);
}

/* overrelax.cu:188-269 (relax_su2 + cabibbo_marinari_relax) */
static void m_cm_relax(m3 g, m3 star) {
  m3 X;
  for (int a = 0; a < 2; a++)
    for (int b = a + 1; b < 3; b++) {
      double w0, w1, w2, w3, det;
      m_mul(X, g, star);
      w0 = X[a][a].re + X[b][b].re;        /* a.x */
      w3 = -X[a][a].im + X[b][b].im;       /* b.y */
      w1 = -X[a][b].im - X[b][a].im;       /* a.y */
      w2 = -X[a][b].re + X[b][a].re;       /* b.x */
      det = w0 * w0 + w1 * w1 + w2 * w2 + w3 * w3;
      det = 1.0 / sqrt(det);
      w0 = w0 * det; w1 = w1 * det; w2 = w2 * det; w3 = w3 * det;
      for (int c = 0; c < 3; c++) {
        cpx d1 = c_mul(c_mk(w0, w3), g[a][c]);
        cpx d2 = c_mul(c_mk(w2, w1), g[b][c]);
        cpx d = c_add(d1, d2);
        cpx e2 = c_mul(c_mk(-w2, w1), g[a][c]);
        cpx e3 = c_mul(c_mk(w0, -w3), g[b][c]);
        g[b][c] = c_add(e2, e3);
        g[a][c] = d;
      }
    }
}

/* overrelax.cu:149-177 (overrelax): g = normalize(sum_{i<N} C(w,i) R^i) g,
 * R = gn g^+ - 1 */
static void m_overrelax(m3 g, m3 gn, double w, int N) {
  m3 one, R, a, Rw, help;
  m_unit(one);
  m_mul_d(R, gn, g);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = c_sub(R[i][j], one[i][j]);
  m_copy(Rw, one);
  m_copy(a, one);
  for (int i = 1; i < N; i++) {
    m_mul(help, a, R);
    m_copy(a, help);
    double f = (1.0 + w - i) / i;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) { a[p][q].re = a[p][q].re * f; a[p][q].im = a[p][q].im * f; }
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) Rw[p][q] = c_add(Rw[p][q], a[p][q]);
  }
  m_normalize(Rw);
  m_mul(help, Rw, g);
  m_copy(g, help);
}

/* The 3-term series of m_overrelax in the short form of the polish sweeps
 * on packed and fixed-point links (the library's overrelax_apply_short):
 * rows 0, 1 of g + c1 (h - g) + c1 c2 (h g^+ h - 2 h + g), which equals
 * (1 + c1 R + c1 c2 R^2) g with R = h g^+ - 1 when g is unitary (the caller
 * normalises).  c1, c2 are formed as overrelax.cu:167 forms them. */
static void m_overrelax_short(m3 g, m3 h, double w) {
  const double c1 = (1.0 + w - 1) / 1, c2 = (1.0 + w - 2) / 2;
  const double cw = c1 * c2, cg = 1.0 - c1 + cw, ch = c1 - 2.0 * cw;
  cpx t[2][3], u[2][3];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = c_mul(h[i][0], c_conj(g[j][0]));
      s = c_add(s, c_mul(h[i][1], c_conj(g[j][1])));
      t[i][j] = c_add(s, c_mul(h[i][2], c_conj(g[j][2])));
    }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = c_mul(t[i][0], h[0][j]);
      s = c_add(s, c_mul(t[i][1], h[1][j]));
      u[i][j] = c_add(s, c_mul(t[i][2], h[2][j]));
    }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++) {
      g[i][j].re = cg * g[i][j].re + ch * h[i][j].re + cw * u[i][j].re;
      g[i][j].im = cg * g[i][j].im + ch * h[i][j].im + cw * u[i][j].im;
    }
}

/* overrelax.cu:451-473 (dev_vectorpotential): U -> (i/2)(U^+ - U), traceless */
static void m_vecpot(m3 U) {
  m3 t;
  cpx ih = c_mk(0.0, 0.5), tr;
  m_dagger(t, U);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) t[i][j] = c_sub(t[i][j], U[i][j]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) U[i][j] = c_mul(ih, t[i][j]);
  tr = m_trace(U);
  tr.re = tr.re / 3.0;
  tr.im = tr.im / 3.0;
  for (int i = 0; i < 3; i++) U[i][i] = c_sub(U[i][i], tr);
}

/* ======================= geometry ======================= */
void orc_init_nn(int LX, int LY, int LZ, int LT, int *nn) {
  for (int t = 0; t < LT; t++)
    for (int z = 0; z < LZ; z++)
      for (int y = 0; y < LY; y++)
        for (int x = 0; x < LX; x++) {
          long p = x + (long)LX * (y + (long)LY * (z + (long)LZ * t));
          nn[8 * p + 3] = x + LX * (y + LY * (z + LZ * ((t + 1) % LT)));
          nn[8 * p + 2] = x + LX * (y + LY * ((z + 1) % LZ + LZ * t));
          nn[8 * p + 1] = x + LX * ((y + 1) % LY + LY * (z + LZ * t));
          nn[8 * p + 0] = (x + 1) % LX + LX * (y + LY * (z + LZ * t));
          nn[8 * p + 7] = x + LX * (y + LY * (z + LZ * ((t + LT - 1) % LT)));
          nn[8 * p + 6] = x + LX * (y + LY * ((z + LZ - 1) % LZ + LZ * t));
          nn[8 * p + 5] = x + LX * ((y + LY - 1) % LY + LY * (z + LZ * t));
          nn[8 * p + 4] = (x + LX - 1) % LX + LX * (y + LY * (z + LZ * t));
        }
}

void orc_init_eo(int LX, int LY, int LZ, int LT, int *eo_even, int *eo_odd) {
  int ne = 0, no = 0;
  for (int t = 0; t < LT; t++)
    for (int z = 0; z < LZ; z++)
      for (int y = 0; y < LY; y++)
        for (int x = 0; x < LX; x++) {
          int p = x + LX * (y + LY * (z + LZ * t));
          if (((x + y + z + t) % 2) == 0) eo_even[ne++] = p;
          else eo_odd[no++] = p;
        }
}

/* ======================= Lehmer RNG ======================= */
#define LEH_MODULUS 2147483647L
#define LEH_MULTIPLIER 48271L
#define LEH_STREAMS 256
#define LEH_A256 22925L
#define LEH_DEFAULT 123456789L
static long leh_seed[LEH_STREAMS] = {LEH_DEFAULT};
static int leh_stream = 0;
static int leh_initialized = 0;

double orc_lehmer_random(void) {
  const long Q = LEH_MODULUS / LEH_MULTIPLIER;
  const long R = LEH_MODULUS % LEH_MULTIPLIER;
  long t = LEH_MULTIPLIER * (leh_seed[leh_stream] % Q) - R * (leh_seed[leh_stream] / Q);
  leh_seed[leh_stream] = (t > 0) ? t : t + LEH_MODULUS;
  return (double)leh_seed[leh_stream] / LEH_MODULUS;
}
void orc_lehmer_put_seed(long x) {
  if (x > 0) x = x % LEH_MODULUS;
  leh_seed[leh_stream] = x;
}
long orc_lehmer_get_seed(void) { return leh_seed[leh_stream]; }
void orc_lehmer_plant_seeds(long x) {
  const long Q = LEH_MODULUS / LEH_A256;
  const long R = LEH_MODULUS % LEH_A256;
  int s = leh_stream;
  leh_initialized = 1;
  leh_stream = 0;
  orc_lehmer_put_seed(x);
  leh_stream = s;
  for (int j = 1; j < LEH_STREAMS; j++) {
    long v = LEH_A256 * (leh_seed[j - 1] % Q) - R * (leh_seed[j - 1] / Q);
    leh_seed[j] = (v > 0) ? v : v + LEH_MODULUS;
  }
}
void orc_lehmer_select_stream(int i) {
  leh_stream = ((unsigned int)i) % LEH_STREAMS;
  if (!leh_initialized && leh_stream != 0) orc_lehmer_plant_seeds(LEH_DEFAULT);
}

static void random_su3(double *m) {
  double help, norm, len;
  cpx r0[3], r1[3], proj, h1, h2;
  help = orc_lehmer_random(); r0[0].re = help; norm = help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r0[0].im = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r0[1].re = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r0[1].im = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r0[2].re = help; norm += help * help;
  r0[2].im = sqrt((1.0 - norm));
  norm = 0.0;
  help = orc_lehmer_random(); r1[0].re = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r1[0].im = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r1[1].re = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r1[1].im = help; norm += help * help;
  help = (1.0 - norm) * orc_lehmer_random(); r1[2].re = help; norm += help * help;
  r1[2].im = sqrt((1.0 - norm));
  proj = c_mk(0.0, 0.0);
  for (int j = 0; j < 3; j++) proj = c_add(proj, c_mul(r1[j], c_conj(r0[j])));
  for (int j = 0; j < 3; j++) r1[j] = c_sub(r1[j], c_mul(proj, r0[j]));
  len = r1[0].re * r1[0].re + r1[0].im * r1[0].im + r1[1].re * r1[1].re +
        r1[1].im * r1[1].im + r1[2].re * r1[2].re + r1[2].im * r1[2].im;
  len = 1.0 / sqrt(len);
  for (int j = 0; j < 3; j++) { r1[j].re = r1[j].re * len; r1[j].im = r1[j].im * len; }
  for (int j = 0; j < 3; j++) {
    m[2 * j] = r0[j].re; m[2 * j + 1] = r0[j].im;
    m[6 + 2 * j] = r1[j].re; m[6 + 2 * j + 1] = r1[j].im;
  }
  /* su3manip.c:617-637: row 2 = (r0 x r1) with the imaginary part negated */
  h1 = c_mul(r0[1], r1[2]); h2 = c_mul(r0[2], r1[1]); h1 = c_sub(h1, h2);
  m[12] = h1.re; m[13] = -h1.im;
  h1 = c_mul(r0[2], r1[0]); h2 = c_mul(r0[0], r1[2]); h1 = c_sub(h1, h2);
  m[14] = h1.re; m[15] = -h1.im;
  h1 = c_mul(r0[0], r1[1]); h2 = c_mul(r0[1], r1[0]); h1 = c_sub(h1, h2);
  m[16] = h1.re; m[17] = -h1.im;
}

void orc_random_init_trafo(double *g, int V) {
  for (long i = 0; i < V; i++) random_su3(g + 18 * i);
}
void orc_random_init_gauge(double *u, int V) {
  for (long i = 0; i < 4L * V; i++) random_su3(u + 18 * i);
}

/* ======================= conversions ======================= */
void orc_su3_to_2v(const double *su3, double *two_v, long n) {
  for (long i = 0; i < n; i++) memcpy(two_v + 12 * i, su3 + 18 * i, 12 * sizeof(double));
}
void orc_2v_to_su3(const double *two_v, double *su3, long n) {
  for (long i = 0; i < n; i++) {
    m3 m;
    m_load(m, two_v + 12 * i);
    memcpy(su3 + 18 * i, m, sizeof(m3));
  }
}

/* ======================= sweep ======================= */
static void site_update(const double *u2v, double *g2v, const int *nn, int pos,
                        int ndir, double omega, int nterms) {
  m3 star, A, B, gx, h;
  m_zero(star);
  for (int mu = 0; mu < ndir; mu++) {
    int hop = nn[8 * pos + mu];
    m_load(A, u2v + 12 * (4L * pos + mu));
    m_load_dagger(B, g2v + 12L * hop);
    m_addmul(star, A, B);
    hop = nn[8 * pos + 4 + mu];
    m_load_dagger(A, u2v + 12 * (4L * hop + mu));
    m_load_dagger(B, g2v + 12L * hop);
    m_addmul(star, A, B);
  }
  m_load(gx, g2v + 12L * pos);
  m_copy(h, gx);
  m_cm_relax(h, star);
  /* nterms < 0: the 3-term series in the short form of the library's
   * polish sweeps (m_overrelax_short), which takes h as the relaxation
   * leaves it; the reference normalises h first (overrelax.cu:413) */
  if (nterms < 0) {
    m_overrelax_short(gx, h, omega);
  } else {
    m_normalize(h);
    m_overrelax(gx, h, omega, nterms);
  }
  m_normalize(gx);
  m_store(g2v + 12L * pos, gx);
}

void orc_or_halfsweep(const double *u2v, double *g2v, const int *nn,
                      const int *sites, int nsites, int ndir, double omega,
                      int nterms, const unsigned char *active, int slice_volume) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nsites; i++) {
    int pos = sites[i];
    if (active && !active[pos / slice_volume]) continue;
    site_update(u2v, g2v, nn, pos, ndir, omega, nterms);
  }
}

/* per-site F and Re tr(Delta Delta^+): overrelax.cu:498-586 */
static void site_functional(const double *u2v, const double *g2v, const int *nn,
                            int pos, int ndir, double *Floc, double *dloc) {
  m3 gather, help, fw, bw, U, G;
  double F = 0.0;
  m_zero(gather);
  for (int mu = 0; mu < ndir; mu++) {
    int hop = nn[8 * pos + mu];
    m_load(U, u2v + 12 * (4L * pos + mu));
    m_load(G, g2v + 12L * pos);
    m_mul(help, G, U);
    m_load_dagger(G, g2v + 12L * hop);
    m_mul(fw, help, G);
    F += m_retrace(fw);
    hop = nn[8 * pos + 4 + mu];
    m_load(U, u2v + 12 * (4L * hop + mu));
    m_load(G, g2v + 12L * hop);
    m_mul(help, G, U);
    m_load_dagger(G, g2v + 12L * pos);
    m_mul(bw, help, G);
    m_vecpot(fw);
    m_vecpot(bw);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) bw[i][j] = c_sub(bw[i][j], fw[i][j]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) gather[i][j] = c_add(gather[i][j], bw[i][j]);
  }
  m_mul_d(help, gather, gather);
  *Floc = F;
  *dloc = 0.0 + m_retrace(help);
}

double orc_functional(const double *u2v, const double *g2v, const int *nn,
                      int V, int ndir, double *Fout, double *dout) {
  int nblk = (V + 127) / 128;
  double *pf = (double *)malloc(sizeof(double) * nblk * 3);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nblk; b++) {
    double f = 0.0, d = 0.0, mx = 0.0;
    for (int t = 0; t < 128; t++) {
      int pos = b * 128 + t;
      if (pos >= V) break;
      double fl, dl;
      site_functional(u2v, g2v, nn, pos, ndir, &fl, &dl);
      f += fl;
      d += dl;
      if (dl > mx) mx = dl;
    }
    pf[3 * b] = f; pf[3 * b + 1] = d; pf[3 * b + 2] = mx;
  }
  double F = 0.0, D = 0.0, M = 0.0;
  for (int b = 0; b < nblk; b++) {
    F += pf[3 * b];
    D += pf[3 * b + 1];
    if (pf[3 * b + 2] > M) M = pf[3 * b + 2];
  }
  free(pf);
  *Fout = F / ((double)ndir * V);
  *dout = D / (double)V;
  return M;
}

void orc_functional_slices(const double *u2v, const double *g2v, const int *nn,
                           int V, int LT, int ndir, double *F, double *dAdA,
                           double *maxdAdA) {
  int vs = V / LT;
#pragma omp parallel for schedule(static)
  for (int t = 0; t < LT; t++) {
    double f = 0.0, d = 0.0, mx = 0.0;
    for (int s = 0; s < vs; s++) {
      double fl, dl;
      site_functional(u2v, g2v, nn, t * vs + s, ndir, &fl, &dl);
      f += fl;
      d += dl;
      if (dl > mx) mx = dl;
    }
    F[t] = f / ((double)ndir * vs);
    dAdA[t] = d / (double)vs;
    maxdAdA[t] = mx;
  }
}

int orc_overrelax_gauge(const double *u2v, double *g2v, int LX, int LY, int LZ,
                        int LT, int maxit, double eps, int ck, double omega,
                        double *hist, int hist_cap, int *nhist) {
  int V = LX * LY * LZ * LT;
  int *nn = (int *)malloc(sizeof(int) * 8L * V);
  int *ev = (int *)malloc(sizeof(int) * (V / 2));
  int *od = (int *)malloc(sizeof(int) * (V / 2));
  double maxd = 0.0, F = 0.0, D = 0.0;
  int i, nh = 0, ret = -1;
  orc_init_nn(LX, LY, LZ, LT, nn);
  orc_init_eo(LX, LY, LZ, LT, ev, od);
  for (i = 0; i < maxit; i++) {
    orc_or_halfsweep(u2v, g2v, nn, ev, V / 2, 4, omega, 3, NULL, V);
    orc_or_halfsweep(u2v, g2v, nn, od, V / 2, 4, omega, 3, NULL, V);
    if (i % ck == 0) {
      maxd = orc_functional(u2v, g2v, nn, V, 4, &F, &D);
      if (hist && nh < hist_cap) {
        hist[3 * nh] = F; hist[3 * nh + 1] = D; hist[3 * nh + 2] = maxd;
      }
      nh++;
    }
    if (maxd < eps) { ret = i; break; }
  }
  if (i == maxit) ret = (maxd > eps) ? -1 : i;
  if (nhist) *nhist = nh;
  free(nn); free(ev); free(od);
  return ret;
}

int orc_coulomb_gauge(const double *u2v, double *g2v, int LX, int LY, int LZ,
                      int LT, int maxit, double eps, int ck, double omega,
                      int *iters) {
  int V = LX * LY * LZ * LT, vs = LX * LY * LZ, left = LT;
  int *nn = (int *)malloc(sizeof(int) * 8L * V);
  int *ev = (int *)malloc(sizeof(int) * (V / 2));
  int *od = (int *)malloc(sizeof(int) * (V / 2));
  unsigned char *act = (unsigned char *)malloc(LT);
  double *F = (double *)malloc(sizeof(double) * LT * 3);
  orc_init_nn(LX, LY, LZ, LT, nn);
  orc_init_eo(LX, LY, LZ, LT, ev, od);
  for (int t = 0; t < LT; t++) { act[t] = 1; iters[t] = -1; }
  for (int i = 0; i < maxit && left > 0; i++) {
    orc_or_halfsweep(u2v, g2v, nn, ev, V / 2, 3, omega, 3, act, vs);
    orc_or_halfsweep(u2v, g2v, nn, od, V / 2, 3, omega, 3, act, vs);
    if (i % ck == 0) {
      orc_functional_slices(u2v, g2v, nn, V, LT, 3, F, F + LT, F + 2 * LT);
      for (int t = 0; t < LT; t++)
        if (act[t] && F[2 * LT + t] < eps) { act[t] = 0; iters[t] = i; left--; }
    }
  }
  free(nn); free(ev); free(od); free(act); free(F);
  return left == 0 ? 0 : -1;
}

double orc_plaquette_2v(const double *u2v, const int *nn, int LX, int LY,
                        int LZ, int LT) {
  int V = LX * LY * LZ * LT;
  double *pt = (double *)calloc(LT, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int t = 0; t < LT; t++) {
    double acc = 0.0;
    for (int z = 0; z < LZ; z++) {
      double mp = 0.0;
      for (int y = 0; y < LY; y++)
        for (int x = 0; x < LX; x++)
          for (int nu = 0; nu < 3; nu++)
            for (int mu = nu + 1; mu < 4; mu++) {
              m3 M1, M2, M3, M4, s1, s2;
              int x0 = x + LX * (y + LY * (z + LZ * t));
              int x1 = nn[8 * x0 + mu], x2 = nn[8 * x0 + nu];
              m_load(M1, u2v + 12 * (4L * x0 + mu));
              m_load(M2, u2v + 12 * (4L * x1 + nu));
              m_load_dagger(M3, u2v + 12 * (4L * x2 + mu));
              m_load_dagger(M4, u2v + 12 * (4L * x0 + nu));
              m_mul(s1, M3, M4);
              m_mul(s2, M2, s1);
              m_mul(s1, M1, s2);
              mp += m_trace(s1).re / 3.0;
            }
      acc += mp;
    }
    pt[t] = acc * (1.0 / (6.0 * V));
  }
  double e = 0.0;
  for (int t = 0; t < LT; t++) e += pt[t];
  free(pt);
  return e;
}

/* ======================= host observables ======================= */
void orc_g_trafo(double *u, const double *g, const int *nn, int V) {
#pragma omp parallel for schedule(static)
  for (int s = 0; s < V; s++)
    for (int mu = 0; mu < 4; mu++) {
      m3 h1, h2, G, U, Gp;
      memcpy(G, g + 18L * s, sizeof(m3));
      memcpy(U, u + 18 * (4L * s + mu), sizeof(m3));
      memcpy(Gp, g + 18L * nn[8 * s + mu], sizeof(m3));
      m_mul(h1, G, U);
      m_dagger(h2, Gp);
      m_mul(U, h1, h2);
      memcpy(u + 18 * (4L * s + mu), U, sizeof(m3));
    }
}

double orc_gauge_functional(const double *u, int V) {
  double h = 0.0;
  for (long k = 0; k < 4L * V; k++) {
    m3 m;
    memcpy(m, u + 18 * k, sizeof(m3));
    h += m_trace(m).re;
  }
  return h / (4.0 * (double)V);
}

/* observables.c:144-165: A = (-i/2)(U - U^+), traceless */
static void host_vecpot(m3 A, m3 U) {
  m3 h1, h2;
  cpx s = c_mk(0.0, -0.5), tr;
  m_dagger(h1, U);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) h2[i][j] = c_sub(U[i][j], h1[i][j]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) A[i][j] = c_mul(s, h2[i][j]);
  tr = m_trace(A);
  tr.re = tr.re / 3.0;
  tr.im = tr.im / 3.0;
  for (int k = 0; k < 3; k++) A[k][k] = c_sub(A[k][k], tr);
}

double orc_dAdA(const double *u, const int *nn, int V, double *maxd) {
  double sum = 0.0, mx = 0.0;
  for (int s = 0; s < V; s++) {
    m3 dA, h1, h2, h4, M1, M2;
    m_zero(dA);
    for (int mu = 0; mu < 4; mu++) {
      int xm = nn[8 * s + 4 + mu];
      memcpy(h1, u + 18 * (4L * xm + mu), sizeof(m3));
      memcpy(h2, u + 18 * (4L * s + mu), sizeof(m3));
      host_vecpot(M1, h1);
      host_vecpot(M2, h2);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dA[i][j] = c_add(dA[i][j], c_sub(M1[i][j], M2[i][j]));
    }
    m_dagger(h1, dA);
    m_mul(h4, dA, h1);
    double tr = m_trace(h4).re;
    sum += tr;
    if (tr > mx) mx = tr;
  }
  if (maxd) *maxd = mx;
  return sum / (double)V;
}

double orc_mean_plaq(const double *u, const int *nn, int V) {
  double mplaq = 0.0;
  for (int x0 = 0; x0 < V; x0++)
    for (int nu = 0; nu < 3; nu++)
      for (int mu = nu + 1; mu < 4; mu++) {
        m3 M1, M2, M3, M4, s1, s2, tmp;
        int x1 = nn[8 * x0 + mu], x2 = nn[8 * x0 + nu];
        memcpy(M1, u + 18 * (4L * x0 + mu), sizeof(m3));
        memcpy(M2, u + 18 * (4L * x1 + nu), sizeof(m3));
        memcpy(tmp, u + 18 * (4L * x2 + mu), sizeof(m3));
        m_dagger(M3, tmp);
        memcpy(tmp, u + 18 * (4L * x0 + nu), sizeof(m3));
        m_dagger(M4, tmp);
        m_mul(s1, M3, M4);
        m_mul(s2, M2, s1);
        m_mul(s1, M1, s2);
        mplaq += m_trace(s1).re / 3.0;
      }
  return mplaq * (1.0 / (6.0 * (double)V));
}

/* ======================= exposed helpers ======================= */
void orc_su3_normalize(double *m18) {
  m3 m;
  memcpy(m, m18, sizeof(m3));
  m_normalize(m);
  memcpy(m18, m, sizeof(m3));
}
void orc_cm_relax(double *g18, const double *star18) {
  m3 g, s;
  memcpy(g, g18, sizeof(m3));
  memcpy(s, star18, sizeof(m3));
  m_cm_relax(g, s);
  memcpy(g18, g, sizeof(m3));
}
void orc_overrelax_step(double *g18, const double *gn18, double w, int N) {
  m3 g, gn;
  memcpy(g, g18, sizeof(m3));
  memcpy(gn, gn18, sizeof(m3));
  m_overrelax(g, gn, w, N);
  memcpy(g18, g, sizeof(m3));
}
void orc_overrelax_step_short(double *g18, const double *h18, double w) {
  m3 g, h;
  memcpy(g, g18, sizeof(m3));
  memcpy(h, h18, sizeof(m3));
  m_overrelax_short(g, h, w);
  m_normalize(g);
  memcpy(g18, g, sizeof(m3));
}
void orc_reconstruct(const double *r12, double *m18) {
  m3 m;
  m_load(m, r12);
  memcpy(m18, m, sizeof(m3));
}
void orc_reconstruct_dagger(const double *r12, double *m18) {
  m3 m;
  m_load_dagger(m, r12);
  memcpy(m18, m, sizeof(m3));
}

/* ======================= Mersenne Twister ======================= */
#define MT_NN 19
#define MT_MM 9
#define MT_UMASK 0xFFFFFFFEu
#define MT_LMASK 0x1u

void orc_mt_generate(orc_mt_params *params, uint32_t *state, int count,
                     int nper, int initialized, float *out, uint32_t *out_words) {
  for (int r = 0; r < count; r++) {
    uint32_t mt[MT_NN];
    orc_mt_params cfg = params[r];
    if (!initialized) {
      mt[0] = cfg.seed;
      for (int s = 1; s < MT_NN; s++)
        mt[s] = 1812433253u * (mt[s - 1] ^ (mt[s - 1] >> 30)) + (uint32_t)s;
    } else {
      for (int s = 0; s < MT_NN; s++) mt[s] = state[(long)r * MT_NN + s];
    }
    int is = 0;
    uint32_t m1 = mt[0];
    for (int k = 0; k < nper; k++) {
      int i1 = is + 1, im = is + MT_MM;
      if (i1 >= MT_NN) i1 -= MT_NN;
      if (im >= MT_NN) im -= MT_NN;
      uint32_t m0 = m1;
      m1 = mt[i1];
      uint32_t mm = mt[im];
      uint32_t x = (m0 & MT_UMASK) | (m1 & MT_LMASK);
      x = mm ^ (x >> 1) ^ ((x & 1u) ? cfg.matrix_a : 0u);
      mt[is] = x;
      is = i1;
      x ^= (x >> 12);
      x ^= (x << 7) & cfg.mask_b;
      x ^= (x << 15) & cfg.mask_c;
      x ^= (x >> 18);
      volatile float fx = (float)x;
      volatile float f1 = fx + 1.0f;
      if (out) out[r + (long)k * count] = f1 / 4294967296.0f;
      if (out_words) out_words[r + (long)k * count] = x;
    }
    params[r].seed = mt[0];
    for (int s = 0; s < MT_NN; s++) state[(long)r * MT_NN + s] = mt[s];
  }
}

uint32_t orc_mt19937_nth(uint32_t seed, int n) {
  /* Matsumoto & Nishimura (1998) MT19937 with the same in-place circular
   * update as above: the generic engine behind the MT607 twisters. */
  enum { N = 624, M = 397 };
  uint32_t mt[N], x = 0;
  mt[0] = seed;
  for (int s = 1; s < N; s++) mt[s] = 1812433253u * (mt[s - 1] ^ (mt[s - 1] >> 30)) + (uint32_t)s;
  int is = 0;
  for (int k = 0; k < n; k++) {
    int i1 = (is + 1) % N, im = (is + M) % N;
    uint32_t y = (mt[is] & 0x80000000u) | (mt[i1] & 0x7fffffffu);
    x = mt[im] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    mt[is] = x;
    is = i1;
    x ^= (x >> 11);
    x ^= (x << 7) & 0x9d2c5680u;
    x ^= (x << 15) & 0xefc60000u;
    x ^= (x >> 18);
  }
  return x;
}

void orc_box_muller_exact(double u1, double u2, double *z1, double *z2) {
  double r = sqrt(-2.0 * log(u1));
  double phi = 2.0 * 3.14159265358979323846 * u2;
  *z1 = r * cos(phi);
  *z2 = r * sin(phi);
}

int orc_mt_nper(int n, int count) {
  int q = (n % count) ? n / count + 1 : n / count;
  return (q % 2) ? q + 1 : q;
}
