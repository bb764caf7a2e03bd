/* ncio.h -- a small reader and writer of netCDF classic files (the
 * published "NetCDF Classic Format Specification": magic "CDF" 1, 32-bit
 * offsets, and "CDF" 2, 64-bit offsets), enough for the model state files
 * the EPIC programs exchange: dimensions (one of them the record
 * dimension), global and variable attributes of type char, int, float and
 * double, and fixed or record variables of type double, float or int.
 * The whole file is held in memory. */
#ifndef EPIC_NCIO_H
#define EPIC_NCIO_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { NCIO_BYTE = 1, NCIO_CHAR = 2, NCIO_SHORT = 3, NCIO_INT = 4, NCIO_FLOAT = 5, NCIO_DOUBLE = 6 };

typedef struct {
  char *name;
  int type;
  size_t n;   /* number of values (characters for NCIO_CHAR) */
  void *val;  /* host-order values: char, int32, float, double */
} ncio_att;

typedef struct {
  char *name;
  size_t len; /* 0: the record dimension */
} ncio_dim;

typedef struct {
  char *name;
  int ndims, dimid[8];
  int natts;
  ncio_att *atts;
  int type;
  /* host-order values of the whole variable, records outermost */
  void *data;
  size_t nval;
  /* placement in the file (filled by the writer and the reader) */
  size_t vsize, begin;
} ncio_var;

typedef struct {
  int version; /* 1 or 2 */
  size_t numrecs;
  int ndims, ngatts, nvars;
  ncio_dim *dims;
  ncio_att *gatts;
  ncio_var *vars;
} ncio_file;

void ncio_init(ncio_file *f, int version);
void ncio_free(ncio_file *f);
/* definitions; each returns the new id or -1 */
int ncio_def_dim(ncio_file *f, const char *name, size_t len);
int ncio_def_var(ncio_file *f, const char *name, int type, int ndims, const int *dimids);
/* attributes of a variable (varid >= 0) or global ones (varid = -1); a
 * second put of a name replaces the first */
int ncio_put_att(ncio_file *f, int varid, const char *name, int type, size_t n, const void *val);
int ncio_put_text(ncio_file *f, int varid, const char *name, const char *text, size_t n);
/* the values of a variable, in host order: all its values (records
 * included, numrecs of them for a record variable) */
int ncio_put_var(ncio_file *f, int varid, const void *val);
/* write the file; 0 or -1 */
int ncio_write(const ncio_file *f, const char *path);
/* read a whole file; 0, or -1 (unreadable) / -2 (not a classic file) */
int ncio_read(ncio_file *f, const char *path);

/* look-ups; NULL or -1 when absent */
const ncio_att *ncio_att_of(const ncio_file *f, int varid, const char *name);
int ncio_var_id(const ncio_file *f, const char *name);
/* number of values of one record (record variables) or of the variable */
size_t ncio_var_count(const ncio_file *f, int varid, int *is_record);

#ifdef __cplusplus
}
#endif
#endif
