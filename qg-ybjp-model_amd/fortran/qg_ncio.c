// Classic netCDF files (CDF-1 "classic" and CDF-2 "64-bit offset") for the
// Fortran driver's r-space dumps and initial conditions, written and read
// without the netCDF library (absent from this toolchain).
//
// The reference writes its dumps through netCDF-Fortran (IO_ncf.f90:15-110
// psi, :116-217 L+A, :376-432 N^2) and reads its initial conditions with
// nf90_open (:223-293, :298-371).  These routines produce and accept the same
// dimensions, variable names, types and values in the classic container,
// which every netCDF reader opens (nf90_open included); the reference's own
// files are netCDF-4 (HDF5) and are rejected with a message naming the
// conversion (nccopy -k 64-bit-offset).
//
// Format (the published classic-format grammar): big-endian throughout;
//   header = 'C' 'D' 'F' version, numrecs, dim_list, gatt_list, var_list
//   dim    = name, length (0 = the record dimension)
//   var    = name, ndims, dimids (slowest first), vatt_list, type, vsize, begin
//   name   = byte count, bytes, zero padding to 4
// begin is 32-bit in version 1 and 64-bit in version 2; fixed-size variables
// are stored contiguously at begin, row-major over their dimensions.
#define _FILE_OFFSET_BITS 64
#define _POSIX_C_SOURCE 200809L
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { MAXN = 64, MAXNAME = 256 };

static __thread char nc_err[512];

const char* qg_nc_error(void) { return nc_err; }

static int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(nc_err, sizeof nc_err, fmt, ap);
  va_end(ap);
  return -1;
}

// ---- names -------------------------------------------------------------------
static int split_csv(const char* csv, char out[][MAXNAME], int max) {
  int n = 0;
  const char* s = csv;
  while (*s) {
    const char* e = strchr(s, ',');
    size_t len = e ? (size_t)(e - s) : strlen(s);
    if (n >= max || len == 0 || len >= MAXNAME) return -1;
    memcpy(out[n], s, len);
    out[n][len] = 0;
    ++n;
    if (!e) break;
    s = e + 1;
  }
  return n;
}

// ---- writing ------------------------------------------------------------------
typedef struct {
  unsigned char* p;
  size_t n, cap;
} Buf;

static void put_bytes(Buf* b, const void* v, size_t n) {
  if (b->n + n > b->cap) {
    b->cap = 2 * (b->n + n) + 256;
    b->p = (unsigned char*)realloc(b->p, b->cap);
  }
  memcpy(b->p + b->n, v, n);
  b->n += n;
}
static void put_u32(Buf* b, uint32_t v) {
  unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8), (unsigned char)v};
  put_bytes(b, c, 4);
}
static void put_u64(Buf* b, uint64_t v) {
  put_u32(b, (uint32_t)(v >> 32));
  put_u32(b, (uint32_t)v);
}
static void put_name(Buf* b, const char* s) {
  static const unsigned char zero[4] = {0, 0, 0, 0};
  const size_t len = strlen(s);
  put_u32(b, (uint32_t)len);
  put_bytes(b, s, len);
  put_bytes(b, zero, (4 - len % 4) % 4);
}

// Header of a CDF-2 file with fixed-size NC_FLOAT variables at the given
// offsets (no attributes).
static void header(Buf* b, char dn[][MAXNAME], const int32_t* dim_len, int nd, char vn[][MAXNAME],
                   const int32_t* var_ndims, const int32_t* var_dims, int nv, const uint64_t* vsize,
                   const uint64_t* begin) {
  b->n = 0;
  put_bytes(b, "CDF\x02", 4);
  put_u32(b, 0);  // numrecs: no record dimension
  put_u32(b, NC_DIMENSION);
  put_u32(b, (uint32_t)nd);
  for (int i = 0; i < nd; ++i) {
    put_name(b, dn[i]);
    put_u32(b, (uint32_t)dim_len[i]);
  }
  put_u32(b, 0);  // no global attributes
  put_u32(b, 0);
  put_u32(b, NC_VARIABLE);
  put_u32(b, (uint32_t)nv);
  int k = 0;
  for (int v = 0; v < nv; ++v) {
    put_name(b, vn[v]);
    put_u32(b, (uint32_t)var_ndims[v]);
    for (int j = 0; j < var_ndims[v]; ++j) put_u32(b, (uint32_t)var_dims[k + j]);
    k += var_ndims[v];
    put_u32(b, 0);  // no variable attributes
    put_u32(b, 0);
    put_u32(b, NC_FLOAT);
    // vsize saturates for a last variable of 4 GiB or more (allowed by the format)
    put_u32(b, vsize[v] >= 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)vsize[v]);
    put_u64(b, begin[v]);
  }
}

// Write a CDF-2 file: dimensions dims_csv ("x,y,z,t") of lengths dim_len,
// NC_FLOAT variables vars_csv, variable v over var_ndims[v] dimensions listed
// (slowest first, as indices into the dimensions) in var_dims, one after the
// other, with values var_data[v] in that row-major order.  Returns 0, or -1
// with qg_nc_error() set.
int qg_nc_write(const char* path, const char* dims_csv, const int32_t* dim_len, int32_t ndims, const char* vars_csv,
                const int32_t* var_ndims, const int32_t* var_dims, const float* const* var_data, int32_t nvars) {
  char dn[MAXN][MAXNAME], vn[MAXN][MAXNAME];
  if (ndims < 1 || ndims > MAXN || split_csv(dims_csv, dn, MAXN) != ndims)
    return fail("%s: %d dimension names expected in \"%s\"", path, (int)ndims, dims_csv);
  if (nvars < 1 || nvars > MAXN || split_csv(vars_csv, vn, MAXN) != nvars)
    return fail("%s: %d variable names expected in \"%s\"", path, (int)nvars, vars_csv);
  for (int i = 0; i < ndims; ++i)
    if (dim_len[i] < 1) return fail("%s: dimension %s has length %d", path, dn[i], (int)dim_len[i]);
  uint64_t vsize[MAXN], begin[MAXN];
  int k = 0;
  for (int v = 0; v < nvars; ++v) {
    if (var_ndims[v] < 0 || var_ndims[v] > ndims) return fail("%s: variable %s has %d dimensions", path, vn[v], (int)var_ndims[v]);
    uint64_t n = 1;
    for (int j = 0; j < var_ndims[v]; ++j) {
      const int d = var_dims[k + j];
      if (d < 0 || d >= ndims) return fail("%s: variable %s: dimension index %d out of range", path, vn[v], d);
      n *= (uint64_t)dim_len[d];
    }
    k += var_ndims[v];
    vsize[v] = 4 * n;
    if (vsize[v] >= 0xFFFFFFFCull && v != nvars - 1)
      return fail("%s: variable %s is 4 GiB or more and is not the last variable", path, vn[v]);
    if (!var_data[v]) return fail("%s: no data for variable %s", path, vn[v]);
  }
  Buf b = {0, 0, 0};
  memset(begin, 0, sizeof begin);
  header(&b, dn, dim_len, ndims, vn, var_ndims, var_dims, nvars, vsize, begin);
  uint64_t off = b.n;
  for (int v = 0; v < nvars; ++v) {
    begin[v] = off;
    off += vsize[v];
  }
  header(&b, dn, dim_len, ndims, vn, var_ndims, var_dims, nvars, vsize, begin);  // same length
  FILE* f = fopen(path, "wb");
  if (!f) {
    free(b.p);
    return fail("%s: cannot create", path);
  }
  int ok = fwrite(b.p, 1, b.n, f) == b.n;
  free(b.p);
  enum { CH = 1 << 16 };
  unsigned char* tmp = (unsigned char*)malloc(4 * CH);
  for (int v = 0; v < nvars && ok; ++v) {
    const uint64_t n = vsize[v] / 4;
    for (uint64_t i0 = 0; i0 < n && ok; i0 += CH) {
      const uint64_t m = n - i0 < CH ? n - i0 : CH;
      for (uint64_t i = 0; i < m; ++i) {
        uint32_t u;
        memcpy(&u, &var_data[v][i0 + i], 4);
        tmp[4 * i] = (unsigned char)(u >> 24);
        tmp[4 * i + 1] = (unsigned char)(u >> 16);
        tmp[4 * i + 2] = (unsigned char)(u >> 8);
        tmp[4 * i + 3] = (unsigned char)u;
      }
      ok = fwrite(tmp, 4, m, f) == m;
    }
  }
  free(tmp);
  if (fclose(f) != 0) ok = 0;
  return ok ? 0 : fail("%s: write failed", path);
}

// ---- reading ------------------------------------------------------------------
typedef struct {
  char name[MAXNAME];
  int32_t ndims, dims[MAXN], type;
  uint64_t begin;
} Var;

static int get(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }
static int get_u32(FILE* f, uint32_t* v) {
  unsigned char c[4];
  if (!get(f, c, 4)) return 0;
  *v = ((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | c[3];
  return 1;
}
static int get_u64(FILE* f, uint64_t* v) {
  uint32_t hi, lo;
  if (!get_u32(f, &hi) || !get_u32(f, &lo)) return 0;
  *v = ((uint64_t)hi << 32) | lo;
  return 1;
}
static int get_name(FILE* f, char* s) {
  uint32_t len;
  if (!get_u32(f, &len) || len >= MAXNAME || !get(f, s, len)) return 0;
  s[len] = 0;
  char pad[4];
  return get(f, pad, (4 - len % 4) % 4);
}
static int type_size(uint32_t t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
  }
}
static int skip_atts(FILE* f) {
  uint32_t tag, n;
  if (!get_u32(f, &tag) || !get_u32(f, &n)) return 0;
  if (tag == 0 && n == 0) return 1;
  if (tag != NC_ATTRIBUTE) return 0;
  for (uint32_t i = 0; i < n; ++i) {
    char name[MAXNAME];
    uint32_t t, m;
    if (!get_name(f, name) || !get_u32(f, &t) || !get_u32(f, &m) || !type_size(t)) return 0;
    const uint64_t bytes = (uint64_t)m * type_size(t);
    if (fseeko(f, (off_t)(bytes + (4 - bytes % 4) % 4), SEEK_CUR) != 0) return 0;
  }
  return 1;
}
static double be_value(const unsigned char* c, int t) {
  switch (t) {
    case NC_BYTE: return (double)(int8_t)c[0];
    case NC_SHORT: return (double)(int16_t)(((unsigned)c[0] << 8) | c[1]);
    case NC_INT: return (double)(int32_t)(((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | c[3]);
    case NC_FLOAT: {
      const uint32_t u = ((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | c[3];
      float x;
      memcpy(&x, &u, 4);
      return x;
    }
    default: {
      uint64_t u = 0;
      for (int i = 0; i < 8; ++i) u = (u << 8) | c[i];
      double x;
      memcpy(&x, &u, 8);
      return x;
    }
  }
}

// Read variable `var` of a CDF-1 or CDF-2 file into out as float (numeric
// types converted as netCDF converts them: rounded to the nearest float).  The
// variable must have exactly the dimensions dims_csv (slowest first) with the
// lengths dim_len; a record dimension counts with the file's record count and
// only its first record is read.  Returns 0, or -1 with qg_nc_error() set.
int qg_nc_read(const char* path, const char* var, const char* dims_csv, const int32_t* dim_len, int32_t ndims,
               float* out) {
  char want[MAXN][MAXNAME];
  if (ndims < 0 || ndims > MAXN || (ndims > 0 && split_csv(dims_csv, want, MAXN) != ndims))
    return fail("%s: %d dimension names expected in \"%s\"", path, (int)ndims, dims_csv);
  FILE* f = fopen(path, "rb");
  if (!f) return fail("%s: cannot open", path);
  unsigned char magic[4];
  char dn[MAXN][MAXNAME];
  uint64_t dl[MAXN];
  uint32_t numrecs, tag, nd, nv;
  int rc = -1;
  Var v;
  if (!get(f, magic, 4)) {
    fail("%s: not a netCDF file", path);
    goto done;
  }
  if (magic[0] == 0x89 && magic[1] == 'H' && magic[2] == 'D' && magic[3] == 'F') {
    fail("%s: netCDF-4 (HDF5) file; convert it to the classic format first (nccopy -k 64-bit-offset)", path);
    goto done;
  }
  if (memcmp(magic, "CDF", 3) != 0 || (magic[3] != 1 && magic[3] != 2)) {
    fail("%s: not a classic (CDF-1/CDF-2) netCDF file", path);
    goto done;
  }
  if (!get_u32(f, &numrecs) || !get_u32(f, &tag) || !get_u32(f, &nd) || (tag != NC_DIMENSION && !(tag == 0 && nd == 0)) ||
      nd > MAXN) {
    fail("%s: bad dimension list", path);
    goto done;
  }
  for (uint32_t i = 0; i < nd; ++i) {
    uint32_t len;
    if (!get_name(f, dn[i]) || !get_u32(f, &len)) {
      fail("%s: bad dimension %u", path, i);
      goto done;
    }
    dl[i] = len ? len : (numrecs == 0xFFFFFFFFu ? 0 : numrecs);
  }
  if (!skip_atts(f)) {
    fail("%s: bad global attributes", path);
    goto done;
  }
  if (!get_u32(f, &tag) || !get_u32(f, &nv) || (tag != NC_VARIABLE && !(tag == 0 && nv == 0))) {
    fail("%s: bad variable list", path);
    goto done;
  }
  for (uint32_t i = 0; i < nv; ++i) {
    uint32_t n, t, vs;
    if (!get_name(f, v.name) || !get_u32(f, &n) || n > MAXN) {
      fail("%s: bad variable %u", path, i);
      goto done;
    }
    v.ndims = (int32_t)n;
    for (uint32_t j = 0; j < n; ++j) {
      uint32_t d;
      if (!get_u32(f, &d) || d >= nd) {
        fail("%s: variable %s: bad dimension id", path, v.name);
        goto done;
      }
      v.dims[j] = (int32_t)d;
    }
    if (!skip_atts(f) || !get_u32(f, &t) || !type_size(t) || !get_u32(f, &vs)) {
      fail("%s: variable %s: bad description", path, v.name);
      goto done;
    }
    v.type = (int32_t)t;
    if (magic[3] == 1) {
      uint32_t b32;
      if (!get_u32(f, &b32)) goto done;
      v.begin = b32;
    } else if (!get_u64(f, &v.begin)) {
      goto done;
    }
    if (strcmp(v.name, var) == 0) break;
    if (i + 1 == nv) {
      fail("%s: no variable %s", path, var);
      goto done;
    }
  }
  if (nv == 0) {
    fail("%s: no variable %s", path, var);
    goto done;
  }
  if (v.type == NC_CHAR) {
    fail("%s: variable %s is text, not numbers", path, var);
    goto done;
  }
  if (v.ndims != ndims) {
    fail("%s: variable %s has %d dimensions, %d expected (%s)", path, var, (int)v.ndims, (int)ndims, dims_csv);
    goto done;
  }
  uint64_t count = 1;
  for (int j = 0; j < ndims; ++j) {
    const int d = v.dims[j];
    if (strcmp(dn[d], want[j]) != 0 || dl[d] != (uint64_t)dim_len[j]) {
      fail("%s: variable %s: dimension %d is %s(%llu), %s(%d) expected", path, var, j, dn[d], (unsigned long long)dl[d],
           want[j], (int)dim_len[j]);
      goto done;
    }
    count *= dl[d];
  }
  if (fseeko(f, (off_t)v.begin, SEEK_SET) != 0) {
    fail("%s: variable %s: bad offset", path, var);
    goto done;
  }
  {
    const int ts = type_size((uint32_t)v.type);
    enum { CH = 1 << 16 };
    unsigned char* tmp = (unsigned char*)malloc((size_t)ts * CH);
    int ok = 1;
    for (uint64_t i0 = 0; i0 < count && ok; i0 += CH) {
      const uint64_t m = count - i0 < CH ? count - i0 : CH;
      ok = fread(tmp, (size_t)ts, m, f) == m;
      for (uint64_t i = 0; ok && i < m; ++i) out[i0 + i] = (float)be_value(tmp + (size_t)ts * i, v.type);
    }
    free(tmp);
    if (!ok) {
      fail("%s: variable %s: file too short", path, var);
      goto done;
    }
  }
  rc = 0;
done:
  fclose(f);
  return rc;
}
