/*
 * ncio.c -- netCDF classic files (CDF-1 and CDF-2) read and written whole,
 * after the published NetCDF Classic Format Specification:
 *
 *   file    = header data
 *   header  = "CDF" version numrecs dim_list gatt_list var_list
 *   list    = ABSENT (two zero words) | tag nelems item...
 *             (tags 0x0A dimensions, 0x0C attributes, 0x0B variables)
 *   name    = nelems bytes, zero-padded to a 4-byte boundary
 *   dim     = name length (0 for the record dimension)
 *   attr    = name type nelems values, zero-padded to 4 bytes
 *   var     = name ndims dimid... vatt_list type vsize begin
 *   data    = every fixed variable at its begin, then numrecs records,
 *             each holding one slice of every record variable in order
 *
 * all integers and values big-endian; begin is 32 bits in CDF-1 and 64 bits
 * in CDF-2.  Fixed variables follow the header in definition order, each
 * padded to 4 bytes; a record holds the record variables' slices padded to
 * 4 bytes, except that a lone record variable is not padded.
 */
#include "ncio.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static size_t type_size(int t) {
  switch (t) {
    case NCIO_BYTE: case NCIO_CHAR: return 1;
    case NCIO_SHORT: return 2;
    case NCIO_INT: case NCIO_FLOAT: return 4;
    case NCIO_DOUBLE: return 8;
    default: return 0;
  }
}
static size_t pad4(size_t n) { return (n + 3) & ~(size_t)3; }
static char *dup_str(const char *s) {
  size_t n = strlen(s);
  char *d = (char *)malloc(n + 1);
  memcpy(d, s, n + 1);
  return d;
}

void ncio_init(ncio_file *f, int version) {
  memset(f, 0, sizeof(*f));
  f->version = (version == 2) ? 2 : 1;
}

static void free_atts(ncio_att *a, int n) {
  for (int i = 0; i < n; i++) {
    free(a[i].name);
    free(a[i].val);
  }
  free(a);
}

void ncio_free(ncio_file *f) {
  for (int i = 0; i < f->ndims; i++) free(f->dims[i].name);
  free(f->dims);
  free_atts(f->gatts, f->ngatts);
  for (int i = 0; i < f->nvars; i++) {
    free(f->vars[i].name);
    free_atts(f->vars[i].atts, f->vars[i].natts);
    free(f->vars[i].data);
  }
  free(f->vars);
  memset(f, 0, sizeof(*f));
}

int ncio_def_dim(ncio_file *f, const char *name, size_t len) {
  f->dims = (ncio_dim *)realloc(f->dims, sizeof(ncio_dim) * (f->ndims + 1));
  f->dims[f->ndims].name = dup_str(name);
  f->dims[f->ndims].len = len;
  return f->ndims++;
}

int ncio_def_var(ncio_file *f, const char *name, int type, int ndims, const int *dimids) {
  if (ndims > 8 || !type_size(type)) return -1;
  for (int d = 0; d < ndims; d++)
    if (dimids[d] < 0 || dimids[d] >= f->ndims || (d > 0 && f->dims[dimids[d]].len == 0)) return -1;
  f->vars = (ncio_var *)realloc(f->vars, sizeof(ncio_var) * (f->nvars + 1));
  ncio_var *v = &f->vars[f->nvars];
  memset(v, 0, sizeof(*v));
  v->name = dup_str(name);
  v->type = type;
  v->ndims = ndims;
  for (int d = 0; d < ndims; d++) v->dimid[d] = dimids[d];
  return f->nvars++;
}

int ncio_put_att(ncio_file *f, int varid, const char *name, int type, size_t n, const void *val) {
  if (!type_size(type) || varid >= f->nvars) return -1;
  ncio_att **list = (varid < 0) ? &f->gatts : &f->vars[varid].atts;
  int *cnt = (varid < 0) ? &f->ngatts : &f->vars[varid].natts;
  int at = -1;
  for (int i = 0; i < *cnt; i++)
    if (!strcmp((*list)[i].name, name)) at = i;
  if (at < 0) {
    *list = (ncio_att *)realloc(*list, sizeof(ncio_att) * (*cnt + 1));
    at = (*cnt)++;
    (*list)[at].name = dup_str(name);
  } else {
    free((*list)[at].val);
  }
  ncio_att *a = &(*list)[at];
  a->type = type;
  a->n = n;
  a->val = malloc(n * type_size(type) + 1);
  if (n) memcpy(a->val, val, n * type_size(type));
  return 0;
}

int ncio_put_text(ncio_file *f, int varid, const char *name, const char *text, size_t n) {
  return ncio_put_att(f, varid, name, NCIO_CHAR, n, text);
}

size_t ncio_var_count(const ncio_file *f, int varid, int *is_record) {
  const ncio_var *v = &f->vars[varid];
  size_t n = 1;
  int rec = 0;
  for (int d = 0; d < v->ndims; d++) {
    size_t len = f->dims[v->dimid[d]].len;
    if (d == 0 && len == 0) rec = 1;
    else n *= len;
  }
  if (is_record) *is_record = rec;
  return n;
}

int ncio_put_var(ncio_file *f, int varid, const void *val) {
  if (varid < 0 || varid >= f->nvars) return -1;
  int rec;
  ncio_var *v = &f->vars[varid];
  size_t n = ncio_var_count(f, varid, &rec) * (rec ? f->numrecs : 1);
  free(v->data);
  v->data = malloc(n * type_size(v->type) + 1);
  if (n) memcpy(v->data, val, n * type_size(v->type));
  v->nval = n;
  return 0;
}

const ncio_att *ncio_att_of(const ncio_file *f, int varid, const char *name) {
  const ncio_att *list = (varid < 0) ? f->gatts : f->vars[varid].atts;
  int cnt = (varid < 0) ? f->ngatts : f->vars[varid].natts;
  for (int i = 0; i < cnt; i++)
    if (!strcmp(list[i].name, name)) return &list[i];
  return NULL;
}

int ncio_var_id(const ncio_file *f, const char *name) {
  for (int i = 0; i < f->nvars; i++)
    if (!strcmp(f->vars[i].name, name)) return i;
  return -1;
}

/* ---------------------------------------------------------------- writing */
typedef struct {
  unsigned char *b;
  size_t n, cap;
} buf_t;

static void put_bytes(buf_t *w, const void *p, size_t n) {
  if (w->n + n > w->cap) {
    size_t c = w->cap ? w->cap : 4096;
    while (c < w->n + n) c *= 2;
    w->b = (unsigned char *)realloc(w->b, c);
    w->cap = c;
  }
  memcpy(w->b + w->n, p, n);
  w->n += n;
}
static void put_u32(buf_t *w, uint32_t v) {
  unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8), (unsigned char)v};
  put_bytes(w, c, 4);
}
static void put_u64(buf_t *w, uint64_t v) {
  put_u32(w, (uint32_t)(v >> 32));
  put_u32(w, (uint32_t)v);
}
static void put_pad(buf_t *w) {
  static const unsigned char z[4] = {0, 0, 0, 0};
  put_bytes(w, z, pad4(w->n) - w->n);
}
static void put_name(buf_t *w, const char *s) {
  size_t n = strlen(s);
  put_u32(w, (uint32_t)n);
  put_bytes(w, s, n);
  put_pad(w);
}
/* n values of type t from host order to big-endian */
static void put_values(buf_t *w, int t, const void *val, size_t n) {
  const size_t sz = type_size(t);
  const unsigned char *p = (const unsigned char *)val;
  unsigned char tmp[8];
  for (size_t i = 0; i < n; i++) {
    for (size_t k = 0; k < sz; k++) tmp[k] = p[i * sz + (sz - 1 - k)];
    put_bytes(w, tmp, sz);
  }
}
static void put_atts(buf_t *w, const ncio_att *a, int n) {
  if (n == 0) {
    put_u32(w, 0);
    put_u32(w, 0);
    return;
  }
  put_u32(w, 0x0C);
  put_u32(w, (uint32_t)n);
  for (int i = 0; i < n; i++) {
    put_name(w, a[i].name);
    put_u32(w, (uint32_t)a[i].type);
    put_u32(w, (uint32_t)a[i].n);
    put_values(w, a[i].type, a[i].val, a[i].n);
    put_pad(w);
  }
}

/* the header with the given begins; its length is the same whatever they are */
static void put_header(buf_t *w, const ncio_file *f) {
  const unsigned char magic[4] = {'C', 'D', 'F', (unsigned char)f->version};
  put_bytes(w, magic, 4);
  put_u32(w, (uint32_t)f->numrecs);
  if (f->ndims == 0) {
    put_u32(w, 0);
    put_u32(w, 0);
  } else {
    put_u32(w, 0x0A);
    put_u32(w, (uint32_t)f->ndims);
    for (int i = 0; i < f->ndims; i++) {
      put_name(w, f->dims[i].name);
      put_u32(w, (uint32_t)f->dims[i].len);
    }
  }
  put_atts(w, f->gatts, f->ngatts);
  if (f->nvars == 0) {
    put_u32(w, 0);
    put_u32(w, 0);
    return;
  }
  put_u32(w, 0x0B);
  put_u32(w, (uint32_t)f->nvars);
  for (int i = 0; i < f->nvars; i++) {
    const ncio_var *v = &f->vars[i];
    put_name(w, v->name);
    put_u32(w, (uint32_t)v->ndims);
    for (int d = 0; d < v->ndims; d++) put_u32(w, (uint32_t)v->dimid[d]);
    put_atts(w, v->atts, v->natts);
    put_u32(w, (uint32_t)v->type);
    put_u32(w, (uint32_t)v->vsize);
    if (f->version == 2) put_u64(w, (uint64_t)v->begin);
    else put_u32(w, (uint32_t)v->begin);
  }
}

int ncio_write(const ncio_file *fc, const char *path) {
  ncio_file f = *fc;  // the placement fields of a copy of the variable list
  ncio_var *vars = (ncio_var *)malloc(sizeof(ncio_var) * (f.nvars ? f.nvars : 1));
  if (f.nvars) memcpy(vars, fc->vars, sizeof(ncio_var) * f.nvars);
  f.vars = vars;
  int nrec = 0;
  for (int i = 0; i < f.nvars; i++) {
    int rec;
    size_t n = ncio_var_count(&f, i, &rec);
    vars[i].vsize = pad4(n * type_size(vars[i].type));
    nrec += rec;
    if (vars[i].nval != n * (rec ? f.numrecs : 1) && vars[i].nval != 0) {
      free(vars);
      return -1;
    }
  }
  buf_t h = {0};
  put_header(&h, &f);
  size_t off = h.n, recsize = 0;
  for (int i = 0; i < f.nvars; i++) {
    int rec;
    ncio_var_count(&f, i, &rec);
    if (!rec) {
      vars[i].begin = off;
      off += vars[i].vsize;
    }
  }
  const size_t rec0 = off;
  for (int i = 0; i < f.nvars; i++) {
    int rec;
    size_t n = ncio_var_count(&f, i, &rec);
    if (rec) {
      vars[i].begin = rec0 + recsize;
      recsize += (nrec == 1) ? n * type_size(vars[i].type) : vars[i].vsize;
    }
  }
  if (f.version == 1 && rec0 + recsize * f.numrecs > 0x7fffffffu) {
    free(vars);
    free(h.b);
    return -1;
  }
  buf_t w = {0};
  put_header(&w, &f);
  free(h.b);
  for (int i = 0; i < f.nvars; i++) {
    int rec;
    size_t n = ncio_var_count(&f, i, &rec);
    if (rec) continue;
    if (vars[i].nval) put_values(&w, vars[i].type, vars[i].data, n);
    else {
      unsigned char *z = (unsigned char *)calloc(n * type_size(vars[i].type) + 1, 1);
      put_bytes(&w, z, n * type_size(vars[i].type));
      free(z);
    }
    put_pad(&w);
  }
  for (size_t r = 0; r < f.numrecs; r++)
    for (int i = 0; i < f.nvars; i++) {
      int rec;
      size_t n = ncio_var_count(&f, i, &rec);
      if (!rec) continue;
      const size_t sz = type_size(vars[i].type);
      if (vars[i].nval) put_values(&w, vars[i].type, (const unsigned char *)vars[i].data + r * n * sz, n);
      else {
        unsigned char *z = (unsigned char *)calloc(n * sz + 1, 1);
        put_bytes(&w, z, n * sz);
        free(z);
      }
      if (nrec > 1) put_pad(&w);
    }
  free(vars);
  FILE *fo = fopen(path, "wb");
  int rc = -1;
  if (fo) {
    rc = (fwrite(w.b, 1, w.n, fo) == w.n) ? 0 : -1;
    if (fclose(fo)) rc = -1;
  }
  free(w.b);
  return rc;
}

/* ---------------------------------------------------------------- reading */
typedef struct {
  const unsigned char *b;
  size_t n, at;
  int bad;
} rd_t;

static uint32_t get_u32(rd_t *r) {
  if (r->at + 4 > r->n) {
    r->bad = 1;
    return 0;
  }
  const unsigned char *p = r->b + r->at;
  r->at += 4;
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint64_t get_u64(rd_t *r) {
  uint64_t hi = get_u32(r);
  return (hi << 32) | get_u32(r);
}
static char *get_name(rd_t *r) {
  size_t n = get_u32(r);
  if (r->bad || r->at + n > r->n) {
    r->bad = 1;
    return dup_str("");
  }
  char *s = (char *)malloc(n + 1);
  memcpy(s, r->b + r->at, n);
  s[n] = 0;
  r->at += pad4(n);
  return s;
}
static void get_values(rd_t *r, int t, void *dst, size_t n, size_t at) {
  const size_t sz = type_size(t);
  if (at + n * sz > r->n || at + n * sz < at) {
    r->bad = 1;
    return;
  }
  unsigned char *d = (unsigned char *)dst;
  const unsigned char *p = r->b + at;
  for (size_t i = 0; i < n; i++)
    for (size_t k = 0; k < sz; k++) d[i * sz + k] = p[i * sz + (sz - 1 - k)];
}
static int get_atts(rd_t *r, ncio_att **out, int *cnt) {
  uint32_t tag = get_u32(r), n = get_u32(r);
  *out = NULL;
  *cnt = 0;
  if (tag == 0 && n == 0) return 0;
  if (tag != 0x0C || n > 100000) return -1;
  *out = (ncio_att *)calloc(n, sizeof(ncio_att));
  *cnt = (int)n;
  for (uint32_t i = 0; i < n && !r->bad; i++) {
    ncio_att *a = &(*out)[i];
    a->name = get_name(r);
    a->type = (int)get_u32(r);
    a->n = get_u32(r);
    const size_t sz = type_size(a->type);
    if (!sz || a->n > r->n) return -1;
    a->val = calloc(a->n * sz + 1, 1);
    get_values(r, a->type, a->val, a->n, r->at);
    r->at += pad4(a->n * sz);
  }
  return r->bad ? -1 : 0;
}

int ncio_read(ncio_file *f, const char *path) {
  memset(f, 0, sizeof(*f));
  FILE *fi = fopen(path, "rb");
  if (!fi) return -1;
  fseek(fi, 0, SEEK_END);
  long len = ftell(fi);
  fseek(fi, 0, SEEK_SET);
  unsigned char *b = (unsigned char *)malloc(len > 0 ? (size_t)len : 1);
  if (len <= 0 || fread(b, 1, (size_t)len, fi) != (size_t)len) {
    fclose(fi);
    free(b);
    return -1;
  }
  fclose(fi);
  rd_t r = {b, (size_t)len, 0, 0};
  int rc = -2;
  if (len < 8 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2)) goto done;
  f->version = b[3];
  r.at = 4;
  {
    uint32_t nr = get_u32(&r);
    if (nr == 0xffffffffu) goto done;  // streaming files are not read here
    f->numrecs = nr;
  }
  {
    uint32_t tag = get_u32(&r), n = get_u32(&r);
    if (!(tag == 0 && n == 0)) {
      if (tag != 0x0A || n > 100000) goto done;
      f->dims = (ncio_dim *)calloc(n, sizeof(ncio_dim));
      f->ndims = (int)n;
      for (uint32_t i = 0; i < n; i++) {
        f->dims[i].name = get_name(&r);
        f->dims[i].len = get_u32(&r);
      }
    }
  }
  if (get_atts(&r, &f->gatts, &f->ngatts)) goto done;
  {
    uint32_t tag = get_u32(&r), n = get_u32(&r);
    if (!(tag == 0 && n == 0)) {
      if (tag != 0x0B || n > 100000) goto done;
      f->vars = (ncio_var *)calloc(n, sizeof(ncio_var));
      f->nvars = (int)n;
      for (uint32_t i = 0; i < n && !r.bad; i++) {
        ncio_var *v = &f->vars[i];
        v->name = get_name(&r);
        v->ndims = (int)get_u32(&r);
        if (v->ndims > 8 || v->ndims < 0) goto done;
        for (int d = 0; d < v->ndims; d++) {
          v->dimid[d] = (int)get_u32(&r);
          if (v->dimid[d] < 0 || v->dimid[d] >= f->ndims) goto done;
        }
        if (get_atts(&r, &v->atts, &v->natts)) goto done;
        v->type = (int)get_u32(&r);
        if (!type_size(v->type)) goto done;
        v->vsize = get_u32(&r);
        v->begin = (f->version == 2) ? (size_t)get_u64(&r) : (size_t)get_u32(&r);
      }
    }
  }
  if (r.bad) goto done;
  {
    // the record size: the record variables' vsizes (a lone one unpadded)
    size_t recsize = 0;
    int nrec = 0;
    for (int i = 0; i < f->nvars; i++) {
      int rec;
      ncio_var_count(f, i, &rec);
      nrec += rec;
    }
    for (int i = 0; i < f->nvars; i++) {
      int rec;
      size_t n = ncio_var_count(f, i, &rec);
      if (rec) recsize += (nrec == 1) ? n * type_size(f->vars[i].type) : f->vars[i].vsize;
    }
    for (int i = 0; i < f->nvars && !r.bad; i++) {
      ncio_var *v = &f->vars[i];
      int rec;
      const size_t n = ncio_var_count(f, i, &rec), sz = type_size(v->type);
      const size_t nr = rec ? f->numrecs : 1;
      v->nval = n * nr;
      v->data = calloc(v->nval * sz + 1, 1);
      for (size_t q = 0; q < nr; q++)
        get_values(&r, v->type, (unsigned char *)v->data + q * n * sz, n, v->begin + q * recsize);
    }
  }
  rc = r.bad ? -2 : 0;
done:
  free(b);
  if (rc) ncio_free(f);
  return rc;
}
