/*
 * gfx_input.c -- the input-file grammar of the reference (read_input.l),
 * parsed by a hand-written line scanner instead of a flex automaton.
 *
 * Accepted (same as read_input.l:75-246):
 *   T = <int>    L = <int>    RandSeed = <int>          (at line start)
 *   BeginORXInit  eps = <flt>  maxit = <int>  ckinterval = <int>  EndORXInit
 *   BeginSAInit   Tmin Tmax expo = <flt>  Nstep ckinterval = <int>  EndSAInit
 *   BeginTHERMInit Nsweep Nsave ckinterval = <int>  beta = <flt>
 *                 start = hot|cold  EndTHERMInit
 *   '#' starts a comment anywhere.
 * A float must contain a '.' (read_input.l:9) and is read through a C float
 * (%f into `float c`, read_input.l:189-190), so eps = 1.0e-12 becomes
 * (double)(float)1e-12.  Anything else on a line is a parsing error.
 *
 * Extension keys (a file without them parses exactly as before):
 *   top level:  MTSeed = <int>
 *   ORX block:  omega = <flt>  precision = fp64|mixed  switch = <flt>
 *               gauge = landau|coulomb
 *               copies = <int>   Gribov copies of every configuration
 *               gaugefiles = <file> <file> ...   (rest of the line, blanks
 *                   or commas between names) configurations fixed in one
 *                   batch together with the -f file
 *   SA block:   micro = <int>  sampler = reference|kp
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cudagfx_host.h"

int randseed = 0;
int orxmaxit = 0, orxcheckinterval = 0, orxflag = 0;
double orxeps = 0.0;
SAPARAM_LIST saparam;
int saflag = 0;
THERMPARAM_LIST thermparam;
int thermflag = 0;
double orxomega = 1.68;
int orxprecision = 0;
double orxswitch = 1e-6;
int orxgauge = 0;
long mtseed = -1;
int samicro = 0;
int sasampler = 0;
int orxcopies = 1;
int norxfiles = 0;
char orxfiles[GFX_MAX_GAUGEFILES][GFX_MAX_PATH];

enum { ST_TOP, ST_ORX, ST_SA, ST_THERM };

static int is_blank(char c) { return c == ' ' || c == '\t'; }

/* length of a {FLT} token at p (read_input.l:9), 0 if none */
static int flt_len(const char *p) {
  const char *q = p;
  if (*q == '+' || *q == '-') q++;
  while (isdigit((unsigned char)*q)) q++;
  if (*q != '.') return 0;
  q++;
  while (isdigit((unsigned char)*q)) q++;
  if (*q == 'e') {
    const char *r = q + 1;
    if (*r == '+' || *r == '-') r++;
    if (isdigit((unsigned char)*r)) {
      while (isdigit((unsigned char)*r)) r++;
      q = r;
    }
  }
  return (int)(q - p);
}

static int int_len(const char *p) {
  const char *q = p;
  while (isdigit((unsigned char)*q)) q++;
  return (int)(q - p);
}

/* "<spaces>key<spaces>=<spaces>" at p; returns chars consumed or 0 */
static int key_eq(const char *p, const char *key) {
  const char *q = p;
  while (is_blank(*q)) q++;
  size_t n = strlen(key);
  if (strncmp(q, key, n) != 0) return 0;
  q += n;
  while (is_blank(*q)) q++;
  if (*q != '=') return 0;
  q++;
  while (is_blank(*q)) q++;
  return (int)(q - p);
}

/* file names from p to the end of the line or a '#', separated by blanks or
 * commas, appended to orxfiles; returns chars consumed, 0 if there is no name
 * or one does not fit */
static int files_len(const char *p) {
  const char *q = p;
  int found = 0;
  while (*q && *q != '#') {
    while (is_blank(*q) || *q == ',') q++;
    if (!*q || *q == '#') break;
    const char *s = q;
    while (*q && *q != '#' && *q != ',' && !is_blank(*q)) q++;
    const size_t len = (size_t)(q - s);
    if (norxfiles >= GFX_MAX_GAUGEFILES || len >= GFX_MAX_PATH) return 0;
    memcpy(orxfiles[norxfiles], s, len);
    orxfiles[norxfiles][len] = '\0';
    norxfiles++;
    found = 1;
  }
  return found ? (int)(q - p) : 0;
}

static float as_float(const char *p, int n) {
  char buf[64];
  if (n > 63) n = 63;
  memcpy(buf, p, n);
  buf[n] = 0;
  float c = 0.f;
  sscanf(buf, "%f", &c);
  return c;
}

static int parse_error(int line, const char *text) {
  printf("Parsing error in line %d\nAborting...!\n", line);
  printf("%s\n", text);
  return 1;
}

/* An unexpected character: the scanner's catch-all rule consumes it and
 * enters its error state (read_input.l:230-237).  Tabs repeat that step;
 * any other character starts the error text and aborts; reaching the end of
 * the line instead silently returns to the initial state (read_input.l:243-246). */
static int unexpected(int line, const char *q, int *state) {
  const char *r = q + 1;
  while (*r == '\t') r++;
  if (*r != '\0') return parse_error(line, r);
  *state = ST_TOP;
  return 0;
}

/* rest of a top-level line after a value: blanks then a comment, nothing,
 * or an unexpected character */
static int top_tail(int line, const char *p, int *state) {
  const char *r = p;
  while (is_blank(*r)) r++;
  if (*r == '\0' && r == p) return 0;
  if (*r == '#') return 0;
  return unexpected(line, p, state);
}

static int parse_block_line(int *state, const char *p, int line) {
  int n, m;
  const char *q;
  while (1) {
    q = p;
    if (*q == '\0') return 0;
    if (*q == '#') return 0;
    {
      const char *r = q;
      while (is_blank(*r)) r++;
      if (*r == '#') return 0;
    }
    if (*state == ST_ORX) {
      if (strncmp(q, "EndORXInit", 10) == 0) {
        orxflag = 1;
        printf("End overrelaxation block in line %d\n\n", line);
        *state = ST_TOP;
        q += 10;
        while (is_blank(*q)) q++;
        return (*q == '\0' || *q == '#') ? 0 : unexpected(line, q, state);
      }
      if ((n = key_eq(q, "eps")) && (m = flt_len(q + n))) {
        float c = as_float(q + n, m);
        orxeps = c;
        printf(" Overrelaxation dAdA epsilon set to %e in line %d\n", c, line);
      } else if ((n = key_eq(q, "maxit")) && (m = int_len(q + n))) {
        orxmaxit = atoi(q + n);
        printf(" Overrelaxation max iterations set to %d line %d\n", orxmaxit, line);
      } else if ((n = key_eq(q, "ckinterval")) && (m = int_len(q + n))) {
        orxcheckinterval = atoi(q + n);
        printf(" Overrelaxation check interval set to %d line %d\n", orxcheckinterval, line);
      } else if ((n = key_eq(q, "omega")) && (m = flt_len(q + n))) {
        orxomega = as_float(q + n, m);
      } else if ((n = key_eq(q, "switch")) && (m = flt_len(q + n))) {
        orxswitch = as_float(q + n, m);
      } else if ((n = key_eq(q, "precision")) && (!strncmp(q + n, "fp64", 4) || !strncmp(q + n, "mixed", 5))) {
        orxprecision = (q[n] == 'm') ? 1 : 0;
        m = (q[n] == 'm') ? 5 : 4;
      } else if ((n = key_eq(q, "gauge")) && (!strncmp(q + n, "landau", 6) || !strncmp(q + n, "coulomb", 7))) {
        orxgauge = (q[n] == 'c') ? 1 : 0;
        m = (q[n] == 'c') ? 7 : 6;
      } else if ((n = key_eq(q, "copies")) && (m = int_len(q + n))) {
        orxcopies = atoi(q + n);
      } else if ((n = key_eq(q, "gaugefiles")) && (m = files_len(q + n))) {
        /* the names run to the end of the line */
      } else {
        return unexpected(line, q, state);
      }
      p = q + n + m;
    } else if (*state == ST_SA) {
      if (strncmp(q, "EndSAInit", 9) == 0) {
        saflag = 1;
        printf("End sim. annealing block in line %d\n\n", line);
        *state = ST_TOP;
        q += 9;
        while (is_blank(*q)) q++;
        return (*q == '\0' || *q == '#') ? 0 : unexpected(line, q, state);
      }
      if ((n = key_eq(q, "Tmin")) && (m = flt_len(q + n))) {
        /* {FLT}+ : several adjacent floats are one token; %f reads the first */
        int k;
        float c = as_float(q + n, m);
        while ((k = flt_len(q + n + m))) m += k;
        saparam.Tmin = c;
        printf(" SA Tmin set to %e in line %d\n", c, line);
      } else if ((n = key_eq(q, "Tmax")) && (m = flt_len(q + n))) {
        int k;
        float c = as_float(q + n, m);
        while ((k = flt_len(q + n + m))) m += k;
        saparam.Tmax = c;
        printf(" SA Tmax set to %e in line %d\n", c, line);
      } else if ((n = key_eq(q, "Nstep")) && (m = int_len(q + n))) {
        saparam.N = atoi(q + n);
        printf(" SA steps set to %d line %d\n", saparam.N, line);
      } else if ((n = key_eq(q, "ckinterval")) && (m = int_len(q + n))) {
        saparam.checkint = atoi(q + n);
        printf(" SA check interval set to %d line %d\n", saparam.checkint, line);
      } else if ((n = key_eq(q, "expo")) && (m = flt_len(q + n))) {
        float c = as_float(q + n, m);
        saparam.expo = c;
        printf(" SA temperature expo set to %e in line %d\n", c, line);
      } else if ((n = key_eq(q, "micro")) && (m = int_len(q + n))) {
        samicro = atoi(q + n);
      } else if ((n = key_eq(q, "sampler")) && (!strncmp(q + n, "reference", 9) || !strncmp(q + n, "kp", 2))) {
        sasampler = (q[n] == 'k') ? 1 : 0;
        m = (q[n] == 'k') ? 2 : 9;
      } else {
        return unexpected(line, q, state);
      }
      p = q + n + m;
    } else { /* ST_THERM */
      if (strncmp(q, "EndTHERMInit", 12) == 0) {
        thermflag = 1;
        printf("End thermalization block in line %d\n\n", line);
        *state = ST_TOP;
        q += 12;
        while (is_blank(*q)) q++;
        return (*q == '\0' || *q == '#') ? 0 : unexpected(line, q, state);
      }
      if ((n = key_eq(q, "Nsweep")) && (m = int_len(q + n))) {
        thermparam.Nsweep = atoi(q + n);
        printf(" Nsweep set to %d line %d\n", thermparam.Nsweep, line);
      } else if ((n = key_eq(q, "Nsave")) && (m = int_len(q + n))) {
        thermparam.Nsave = atoi(q + n);
        printf(" Nsave set to %d line %d\n", thermparam.Nsave, line);
      } else if ((n = key_eq(q, "ckinterval")) && (m = int_len(q + n))) {
        thermparam.checkint = atoi(q + n);
        printf(" Therm. check interval set to %d line %d\n", thermparam.checkint, line);
      } else if ((n = key_eq(q, "beta")) && (m = flt_len(q + n))) {
        float c = as_float(q + n, m);
        thermparam.beta = c;
        printf(" beta set to %e in line %d\n", c, line);
      } else if ((n = key_eq(q, "start")) && !strncmp(q + n, "hot", 3)) {
        thermparam.startcond = 1;
        m = 3;
        printf(" using hot start in line %d\n", line);
      } else if ((n = key_eq(q, "start")) && !strncmp(q + n, "cold", 4)) {
        thermparam.startcond = 0;
        m = 4;
        printf(" using cold start in line %d\n", line);
      } else {
        return unexpected(line, q, state);
      }
      p = q + n + m;
    }
  }
}

static int parse_top_line(int *state, const char *p, int line) {
  int n, m;
  if (*p == '\0' || *p == '#') return 0;
  {
    const char *r = p;
    while (is_blank(*r)) r++;
    if (*r == '#') return 0;
  }
  {
    /* ^T= ^L= ^RandSeed= (read_input.l:75-96): digits are optional, as in
     * the scanner, which then just returns to the initial state */
    const char *keys[4] = {"T", "L", "RandSeed", "MTSeed"};
    for (int k = 0; k < 4; k++) {
      if (p[0] != keys[k][0] || !(n = key_eq(p, keys[k]))) continue;
      m = int_len(p + n);
      if (m) {
        if (k == 0) {
          T = atoi(p + n);
          printf("T = %.*s\n", m, p + n);
        } else if (k == 1) {
          L = atoi(p + n);
          printf("L = %.*s\n", m, p + n);
        } else if (k == 2) {
          randseed = atoi(p + n);
          printf("Random Seed = %.*s\n", m, p + n);
        } else {
          mtseed = atol(p + n);
        }
      }
      return top_tail(line, p + n + m, state);
    }
  }
  const char *blocks[3] = {"BeginORX", "BeginSA", "BeginTHERM"};
  const int states[3] = {ST_ORX, ST_SA, ST_THERM};
  const char *msg[3] = {"Found overrelaxation block in line %d\n", "Found sim. annealing block in line %d\n",
                        "Found thermalization block in line %d\n"};
  for (int b = 0; b < 3; b++) {
    size_t k = strlen(blocks[b]);
    if (strncmp(p, blocks[b], k) == 0) {
      const char *q = p + k;
      if (strncmp(q, "Init", 4) != 0) {
        *state = states[b];
        return (*q == '\0') ? 0 : unexpected(line, q, state);
      }
      q += 4;
      while (is_blank(*q)) q++;
      printf(msg[b], line);
      *state = states[b];
      if (*q == '\0' || *q == '#') return 0;
      return parse_block_line(state, q, line);
    }
  }
  return unexpected(line, p, state);
}

int read_input(char *conf_file) {
  FILE *f = fopen(conf_file, "rt");
  if (!f) return 2;
  char buf[4096];
  int line = 1, state = ST_TOP, rc = 0;
  norxfiles = 0; /* the list is the one of this file */
  while (fgets(buf, sizeof buf, f)) {
    size_t n = strlen(buf);
    while (n > 0 && (buf[n - 1] == '\n' || buf[n - 1] == '\r')) buf[--n] = '\0';
    rc = (state == ST_TOP) ? parse_top_line(&state, buf, line) : parse_block_line(&state, buf, line);
    if (rc) break;
    line++;
  }
  fclose(f);
  if (rc) return 1;
  LX = L;
  LY = L;
  LZ = L;
  VOLUME = LX * LY * LZ * T;
  return 0;
}
