/*
 * gfx_ildg.c -- ILDG gauge files in LIME containers, read and written
 * without c-lime.
 *
 * LIME record (public LIME format, c-lime 1.3.2): a 144-byte big-endian
 * header {uint32 magic 0x456789ab; uint16 version 1; uint16 flags with
 * bit 15 = message begin, bit 14 = message end; uint64 data length;
 * char type[128]} followed by the data, zero-padded to a multiple of 8.
 *
 * ILDG content as the reference reads and writes it (gauge_io.c:245-595):
 *   "ildg-format"      XML with <precision>, <lx> <ly> <lz> <lt>
 *   "xlf-info"         text with "plaquette = x" and "beta = x"
 *   "ildg-binary-data" 4V su3 matrices, big-endian, site-major in
 *                      s = x + LX*(y + LY*(z + LZ*t)), mu-minor, row-major
 * Each record is written as its own message (MB = ME = 1).
 */
#define _FILE_OFFSET_BITS 64
#include <stdint.h>
#include <stdio.h>
#include <sys/types.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cudagfx_host.h"

#define LIME_MAGIC 0x456789abu
#define LIME_HDR 144

static void put_be32(unsigned char *p, uint32_t v) {
  for (int k = 0; k < 4; k++) p[k] = (unsigned char)(v >> (24 - 8 * k));
}
static void put_be16(unsigned char *p, uint16_t v) {
  p[0] = (unsigned char)(v >> 8);
  p[1] = (unsigned char)v;
}
static void put_be64(unsigned char *p, uint64_t v) {
  for (int k = 0; k < 8; k++) p[k] = (unsigned char)(v >> (56 - 8 * k));
}
static uint32_t get_be32(const unsigned char *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint64_t get_be64(const unsigned char *p) {
  uint64_t v = 0;
  for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
  return v;
}

static void swap_bytes(void *buf, size_t n, size_t width) {
  unsigned char *b = (unsigned char *)buf;
  for (size_t i = 0; i < n; i++, b += width)
    for (size_t k = 0; k < width / 2; k++) {
      unsigned char t = b[k];
      b[k] = b[width - 1 - k];
      b[width - 1 - k] = t;
    }
}

static int host_is_little_endian(void) {
  const uint16_t one = 1;
  return *(const unsigned char *)&one == 1;
}

/* content between <tag> and </tag> (spinor_io.c:310-339 behaviour) */
static int xml_entry(const char *s, const char *tag, char *out, size_t cap) {
  char b[256], e[256];
  snprintf(b, sizeof b, "<%s>", tag);
  snprintf(e, sizeof e, "</%s>", tag);
  const char *p = strstr(s, b);
  if (!p) return 0;
  const char *q = strstr(s, e);
  if (!q) return 0;
  p += strlen(b);
  size_t n = (q > p) ? (size_t)(q - p) : 0;
  if (n >= cap) n = cap - 1;
  memcpy(out, p, n);
  out[n] = 0;
  return 1;
}

/* value after "<name> =" in whitespace-separated text (gauge_io.c:119-135) */
static int text_entry(const char *s, const char *name, char *out, size_t cap) {
  char *dup = strdup(s), *save = NULL;
  int found = 0;
  if (!dup) return 0;
  for (char *t = strtok_r(dup, " \n\t", &save); t; t = strtok_r(NULL, " \n\t", &save)) {
    if (strcmp(t, name) == 0) {
      char *eq = strtok_r(NULL, " \n\t", &save);
      char *val = eq ? strtok_r(NULL, " \n\t", &save) : NULL;
      if (val) {
        snprintf(out, cap, "%s", val);
        found = 1;
      }
      break;
    }
  }
  free(dup);
  return found;
}

static int lime_write_record(FILE *f, const char *type, const void *data, uint64_t bytes) {
  unsigned char h[LIME_HDR];
  memset(h, 0, sizeof h);
  put_be32(h, LIME_MAGIC);
  put_be16(h + 4, 1);
  put_be16(h + 6, (uint16_t)((1u << 15) | (1u << 14)));
  put_be64(h + 8, bytes);
  strncpy((char *)h + 16, type, 127);
  if (fwrite(h, 1, LIME_HDR, f) != LIME_HDR) return -1;
  if (data && bytes && fwrite(data, 1, bytes, f) != bytes) return -1;
  static const unsigned char zeros[8] = {0};
  const size_t pad = (size_t)((8 - bytes % 8) % 8);
  if (pad && fwrite(zeros, 1, pad, f) != pad) return -1;
  return 0;
}

/* LIME records are read from files that may be truncated or corrupt: every
 * length is checked against the bytes that are left in the file before it
 * is used, and metadata records (ildg-format, xlf-info) are held in memory
 * only up to LIME_META_MAX bytes. */
#define LIME_META_MAX (1u << 20)

/* size of an open file, or -1 (also for input that cannot seek) */
static int64_t file_size(FILE *f) {
  if (fseeko(f, 0, SEEK_END) != 0) return -1;
  const off_t n = ftello(f);
  if (n < 0 || fseeko(f, 0, SEEK_SET) != 0) return -1;
  return (int64_t)n;
}

/* the next record header: 1 with *len (checked to fit in the rest of the
 * file) and the type, 0 at the end of the file, -1 for a short header, a
 * wrong magic number or data running past the end of the file */
static int lime_next(FILE *f, int64_t fsize, uint64_t *len, char type[129]) {
  unsigned char h[LIME_HDR];
  const size_t got = fread(h, 1, LIME_HDR, f);
  if (got == 0 && feof(f)) return 0;
  if (got != LIME_HDR || get_be32(h) != LIME_MAGIC) return -1;
  const off_t pos = ftello(f);
  if (pos < 0 || (int64_t)pos > fsize) return -1;
  *len = get_be64(h + 8);
  if (*len > (uint64_t)(fsize - (int64_t)pos)) return -1;
  memcpy(type, h + 16, 128);
  type[128] = 0;
  return 1;
}

/* moves past the rest of a record whose data has length len (done bytes of
 * it already read) and its padding to a multiple of 8; the padding of the
 * last record may be cut short by the end of the file */
static int lime_skip(FILE *f, int64_t fsize, uint64_t len, uint64_t done) {
  const off_t pos = ftello(f);
  if (pos < 0 || (int64_t)pos > fsize) return -1;
  const uint64_t left = (uint64_t)(fsize - (int64_t)pos);
  uint64_t skip = len - done + (8 - len % 8) % 8;
  if (skip > left) skip = left;
  return fseeko(f, (off_t)skip, SEEK_CUR) == 0 ? 0 : -1;
}

/* the data of a metadata record as a NUL-terminated string (NULL if it is
 * longer than LIME_META_MAX or cannot be read) */
static char *lime_read_text(FILE *f, uint64_t len) {
  if (len > LIME_META_MAX) return NULL;
  char *msg = (char *)calloc((size_t)len + 1, 1);
  if (!msg) return NULL;
  if (len && fread(msg, 1, (size_t)len, f) != (size_t)len) {
    free(msg);
    return NULL;
  }
  return msg;
}

int read_ildg_dims(char *filename, int dims[4]) {
  FILE *f = fopen(filename, "rb");
  if (!f) return -1;
  const int64_t fsize = file_size(f);
  int rc = -1;
  uint64_t len;
  char type[129];
  while (fsize >= 0 && lime_next(f, fsize, &len, type) == 1) {
    if (!strcmp(type, "ildg-binary-data")) break;
    if (strcmp(type, "ildg-format") != 0 || len > LIME_META_MAX) {
      if (lime_skip(f, fsize, len, 0) != 0) break;
      continue;
    }
    char *msg = lime_read_text(f, len);
    if (msg) {
      const char *tags[4] = {"lx", "ly", "lz", "lt"};
      char v[64];
      rc = 0;
      for (int k = 0; k < 4; k++) {
        dims[k] = 0;
        if (!xml_entry(msg, tags[k], v, sizeof v) || sscanf(v, "%d", &dims[k]) != 1 || dims[k] <= 0)
          rc = -1;
      }
    }
    free(msg);
    break;
  }
  fclose(f);
  return rc;
}

int read_gf_ildg(gfx_su3 *gf, char *filename) {
  FILE *f = fopen(filename, "rb");
  if (!f) {
    fprintf(stderr, "Error opening file %s\n", filename);
    return -1;
  }
  printf("Reading ildg-format:\n");
  printf("Reading file '%s' in ildg-format\n", filename);
  const int64_t fsize = file_size(f);
  if (fsize < 0) {
    fprintf(stderr, "Error in read_gf_ildg: cannot determine the size of %s\n", filename);
    fclose(f);
    return -1;
  }
  int precision = 0, readprecision = 0, readformat = 0, have_data = 0;
  uint64_t data_bytes = 0, len;
  char type[129];
  for (;;) {
    const int st = lime_next(f, fsize, &len, type);
    if (st == 0) break;
    if (st < 0) {
      fprintf(stderr, "limeReaderNextRecord returned error with status = %d!\n", -1);
      break;
    }
    if (!strcmp(type, "ildg-binary-data")) {
      data_bytes = len;
      have_data = 1;
      break;
    }
    if (!strcmp(type, "ildg-format") || !strcmp(type, "xlf-info")) {
      char *msg = lime_read_text(f, len);
      if (!msg) {
        fprintf(stderr, "Error in read_gf_ildg: %s record of %llu bytes cannot be read\n", type,
                (unsigned long long)len);
        break;
      }
      if (lime_skip(f, fsize, len, len) != 0) {
        free(msg);
        break;
      }
      char v[512];
      if (!strcmp(type, "ildg-format")) {
        if (xml_entry(msg, "field", v, sizeof v)) printf("The field is of type: %s\n", v);
        if (xml_entry(msg, "precision", v, sizeof v)) sscanf(v, "%d", &precision);
        readprecision = 1;
        printf("The precision is: %d\n", precision);
        const char *tags[4] = {"lx", "ly", "lz", "lt"};
        const int want[4] = {LX, LY, LZ, T};
        const char *lab[4] = {"LX", "LY", "LZ", "T"};
        for (int k = 0; k < 4; k++) {
          int d = 0;
          if (xml_entry(msg, tags[k], v, sizeof v)) sscanf(v, "%d", &d);
          printf("%s: %d\n", lab[k], d);
          if (d != want[k]) {
            fprintf(stderr, "Error in read_gf_ildg: wrong %s dimension in gauge file. Aborting...\n", lab[k]);
            free(msg);
            fclose(f);
            return 120;
          }
        }
        VOLUME = LX * LY * LZ * T;
        printf("VOLUME: %d\n", VOLUME);
        readformat = 1;
      } else {
        if (text_entry(msg, "plaquette", v, sizeof v)) PLAQ = atof(v);
        printf("plaquette: %e\n", PLAQ);
        if (text_entry(msg, "beta", v, sizeof v)) BETA = atof(v);
        printf("beta: %e\n", BETA);
      }
      free(msg);
    } else if (lime_skip(f, fsize, len, 0) != 0) {
      break;
    }
  }
  if (!have_data || !readformat) {
    if (!readformat) fprintf(stderr, "Unable to read Xml-Format (ildg-format) out of LimeReader\n");
    fclose(f);
    return -1;
  }
  const uint64_t n = (uint64_t)4 * LX * LY * LZ * T * 18;
  const int le = host_is_little_endian();
  if (data_bytes == n * sizeof(float) && (precision == 32 || !readprecision)) {
    float *buf = (float *)malloc(sizeof(float) * 18 * 4096);
    double *out = (double *)gf;
    if (!buf) {
      fclose(f);
      return -1;
    }
    for (uint64_t done = 0; done < n;) {
      const size_t chunk = (size_t)((n - done < 18 * 4096) ? n - done : 18 * 4096);
      if (fread(buf, sizeof(float), chunk, f) != chunk) {
        free(buf);
        fclose(f);
        return 500;
      }
      if (le) swap_bytes(buf, chunk, sizeof(float));
      for (size_t k = 0; k < chunk; k++) out[done + k] = (double)buf[k];
      done += chunk;
    }
    free(buf);
  } else if (data_bytes == n * sizeof(double) && (precision == 64 || !readprecision)) {
    if (fread(gf, sizeof(double), n, f) != n) {
      fclose(f);
      return 500;
    }
    if (le) swap_bytes(gf, n, sizeof(double));
  } else {
    fprintf(stderr, "Error in %s: Xml-Entry 'precision' does not match size of data block. Possible file corruption! Aborting...\n", filename);
    fclose(f);
    return 500;
  }
  fclose(f);
  printf("\n\n");
  return 0;
}

int write_gf_ildg(gfx_su3 *gf, char *filename, const int prec) {
  if (prec != 32 && prec != 64) {
    fprintf(stderr, "Wrong precision! (must be 32 or 64) Aborting...\n");
    return 500;
  }
  FILE *f = fopen(filename, "wb");
  if (!f) {
    fprintf(stderr, "Error opening file %s\n", filename);
    return -1;
  }
  printf("Writing ildg-format:\n");
  printf("Writing to file '%s' in ildg-format\n", filename);
  char msg[2048];
  snprintf(msg, sizeof msg,
           "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqcd.org/ildg\"\n"
           "            xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
           "            xsi:schemaLocation=\"http://www.lqcd.org/ildg filefmt.xsd\">\n"
           "  <version> 1.0 </version>\n  <field> su3gauge </field>\n  <precision> %d </precision>\n"
           "  <lx> %d </lx>\n  <ly> %d </ly>\n  <lz> %d </lz>\n  <lt> %d </lt>\n</ildgFormat>",
           prec, LX, LY, LZ, T);
  int rc = lime_write_record(f, "ildg-format", msg, strlen(msg));
  snprintf(msg, sizeof msg,
           " plaquette = %.6f\ntrajectory nr = \nbeta = %.6f, kappa = , mu = , c2_rec = \n"
           "time = 1216028646\nhmcversion = 4.0.1\nmubar = \nepsilonbar = \ndate = ",
           PLAQ, BETA);
  rc |= lime_write_record(f, "xlf-info", msg, strlen(msg));
  const uint64_t n = (uint64_t)4 * LX * LY * LZ * T * 18;
  const uint64_t bytes = n * (prec == 64 ? 8 : 4);
  /* header, then the data streamed in chunks, then padding */
  unsigned char h[LIME_HDR];
  memset(h, 0, sizeof h);
  put_be32(h, LIME_MAGIC);
  put_be16(h + 4, 1);
  put_be16(h + 6, (uint16_t)((1u << 15) | (1u << 14)));
  put_be64(h + 8, bytes);
  strncpy((char *)h + 16, "ildg-binary-data", 127);
  if (fwrite(h, 1, LIME_HDR, f) != LIME_HDR) rc = -1;
  const int le = host_is_little_endian();
  const double *src = (const double *)gf;
  const size_t CH = 18 * 4096;
  void *buf = malloc(CH * 8);
  if (!buf) rc = -1;
  for (uint64_t done = 0; done < n && rc == 0;) {
    const size_t chunk = (size_t)((n - done < CH) ? n - done : CH);
    if (prec == 64) {
      memcpy(buf, src + done, chunk * 8);
      if (le) swap_bytes(buf, chunk, 8);
      if (fwrite(buf, 8, chunk, f) != chunk) rc = -1;
    } else {
      float *fb = (float *)buf;
      for (size_t k = 0; k < chunk; k++) fb[k] = (float)src[done + k];
      if (le) swap_bytes(buf, chunk, 4);
      if (fwrite(buf, 4, chunk, f) != chunk) rc = -1;
    }
    done += chunk;
  }
  free(buf);
  /* the binary block is a multiple of 8 bytes except in 32-bit with odd n */
  if (rc == 0 && bytes % 8) {
    static const unsigned char zeros[8] = {0};
    if (fwrite(zeros, 1, 8 - bytes % 8, f) != 8 - bytes % 8) rc = -1;
  }
  fflush(f);
  fclose(f);
  printf("\n\n");
  return rc ? -1 : 0;
}
