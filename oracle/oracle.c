/*
 * oracle.c -- CPU restatement of monolish's reference path (see oracle.h).
 * TEST INFRASTRUCTURE ONLY: the product never links or calls this file.
 * Built with -ffp-contract=off so every multiply and add rounds separately,
 * like the reference's -O3 -mavx (no -mfma) build.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int or_num_threads(void) { return omp_get_max_threads(); }

static double g_sor_omega = 1.9;
void or_set_sor_omega(double w) { g_sor_omega = w; }
double or_get_sor_omega(void) { return g_sor_omega; }
static int64_t g_factor_bad_row = -1;
int64_t or_factor_bad_row(void) { return g_factor_bad_row; }

/* ---------------------------------------------------------------------------
 * std::mt19937 (the C++ standard's mersenne_twister_engine<uint32_t, 32, 624,
 * 397, 31, 0x9908b0df, 11, 0xffffffff, 7, 0x9d2c5680, 15, 0xefc60000, 18,
 * 1812433253>).  Used by vector(N,min,max,seed) and util::random_vector,
 * src/utils/vector_constructor.cpp:78-113.
 * ------------------------------------------------------------------------- */
void or_mt_seed(or_mt19937 *s, uint32_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 624; i++) {
    uint32_t prev = s->mt[i - 1];
    s->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
  }
  s->idx = 624;
}

static void or_mt_twist(or_mt19937 *s) {
  for (int i = 0; i < 624; i++) {
    uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
    uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
    if (y & 1u)
      v ^= 0x9908b0dfu;
    s->mt[i] = v;
  }
  s->idx = 0;
}

uint32_t or_mt_next(or_mt19937 *s) {
  if (s->idx >= 624)
    or_mt_twist(s);
  uint32_t y = s->mt[s->idx++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* std::uniform_real_distribution<double>(a, b) driven by mt19937, as
 * libstdc++ implements it: generate_canonical<double, 53> takes two 32-bit
 * draws u1, u2, forms (u1 + u2*2^32) / 2^64 in double (clamped below 1), and
 * the distribution returns u*(b-a) + a. */
double or_mt_uniform(or_mt19937 *s, double a, double b) {
  double u1 = (double)or_mt_next(s);
  double u2 = (double)or_mt_next(s);
  const double r = 4294967296.0; /* 2^32 */
  double sum = u1;
  double t = u2 * r;
  sum = sum + t;
  double u = sum / (r * r);
  if (u >= 1.0)
    u = nextafter(1.0, 0.0);
  double span = b - a;
  double scaled = u * span;
  return scaled + a;
}

void or_random_vector_d(double *v, int64_t n, double min, double max,
                        uint32_t seed) {
  or_mt19937 s;
  or_mt_seed(&s, seed);
  for (int64_t i = 0; i < n; i++)
    v[i] = or_mt_uniform(&s, min, max);
}

/* float vectors draw the double sample and round it to float
 * (vector_constructor.cpp:98-113 with T = float) */
void or_random_vector_s(float *v, int64_t n, float min, float max,
                        uint32_t seed) {
  or_mt19937 s;
  or_mt_seed(&s, seed);
  for (int64_t i = 0; i < n; i++)
    v[i] = (float)or_mt_uniform(&s, (double)min, (double)max);
}

/* ---------------------------------------------------------------------------
 * COO
 * ------------------------------------------------------------------------- */
void or_coo_init(or_coo *m, int64_t rows, int64_t cols) {
  m->rows = rows;
  m->cols = cols;
  m->nnz = 0;
  m->alloc = 0;
  m->row = NULL;
  m->col = NULL;
  m->val = NULL;
}

void or_coo_free(or_coo *m) {
  free(m->row);
  free(m->col);
  free(m->val);
  m->row = m->col = NULL;
  m->val = NULL;
  m->nnz = m->alloc = 0;
}

/* COO::insert -- src/utils/at_insert_sort/at_insert_sort_coo.cpp:29-51:
 * append; capacity grows to 2*alloc+1 when full. */
void or_coo_insert(or_coo *m, int64_t i, int64_t j, double v) {
  if (m->nnz >= m->alloc) {
    int64_t na = 2 * m->alloc + 1;
    m->row = (int32_t *)realloc(m->row, sizeof(int32_t) * (size_t)na);
    m->col = (int32_t *)realloc(m->col, sizeof(int32_t) * (size_t)na);
    m->val = (double *)realloc(m->val, sizeof(double) * (size_t)na);
    m->alloc = na;
  }
  m->row[m->nnz] = (int32_t)i;
  m->col[m->nnz] = (int32_t)j;
  m->val[m->nnz] = v;
  m->nnz++;
}

/* COO::at -- at_insert_sort_coo.cpp:9-24: search from the last entry so the
 * last inserted duplicate is the one returned. */
double or_coo_at(const or_coo *m, int64_t i, int64_t j) {
  for (int64_t k = m->nnz; k > 0; --k)
    if (m->row[k - 1] == (int32_t)i && m->col[k - 1] == (int32_t)j)
      return m->val[k - 1];
  return 0.0;
}

static int coo_key_cmp(const or_coo *m, int64_t a, int64_t b) {
  return (m->row[a] != m->row[b]) ? (m->row[a] - m->row[b])
                                  : (m->col[a] - m->col[b]);
}

static void coo_swap(or_coo *m, int64_t a, int64_t b) {
  int32_t t = m->row[a];
  m->row[a] = m->row[b];
  m->row[b] = t;
  t = m->col[a];
  m->col[a] = m->col[b];
  m->col[b] = t;
  double v = m->val[a];
  m->val[a] = m->val[b];
  m->val[b] = v;
}

/* COO::_q_sort -- at_insert_sort_coo.cpp:54-110.  The same partition scheme
 * (pivot = last element; left scan while key <= pivot, right scan while
 * key >= pivot, swap, then drop the pivot at the meeting point) is kept
 * because the order it leaves duplicates in decides which value the merge
 * keeps.  Sub-ranges are disjoint, so an explicit stack replaces the
 * recursion without changing the result. */
static void coo_qsort(or_coo *m, int64_t lo0, int64_t hi0) {
  int64_t cap = 64, top = 0;
  int64_t *st = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)cap);
  st[top++] = lo0;
  st[top++] = hi0;
  while (top > 0) {
    int64_t hi = st[--top];
    int64_t lo = st[--top];
    if (lo >= hi)
      continue;
    int64_t l = lo, h = hi, p = hi;
    int32_t p1 = m->row[p], p2 = m->col[p];
    double p3 = m->val[p];
    do {
      while (l < h && coo_key_cmp(m, l, p) <= 0)
        l++;
      while (h > l && coo_key_cmp(m, h, p) >= 0)
        h--;
      if (l < h)
        coo_swap(m, l, h);
    } while (l < h);
    m->row[p] = m->row[l];
    m->row[l] = p1;
    m->col[p] = m->col[l];
    m->col[l] = p2;
    m->val[p] = m->val[l];
    m->val[l] = p3;
    if (top + 4 > 2 * cap) {
      cap *= 2;
      st = (int64_t *)realloc(st, sizeof(int64_t) * 2 * (size_t)cap);
    }
    st[top++] = lo;
    st[top++] = l - 1;
    st[top++] = l + 1;
    st[top++] = hi;
  }
  free(st);
}

/* COO::sort -- at_insert_sort_coo.cpp:113-141: sort by (row, col); with
 * merge, duplicates collapse and the value of the last one (in sorted
 * order) is kept.  An empty COO stays empty. */
void or_coo_sort(or_coo *m, int merge) {
  if (m->nnz <= 0)
    return;
  coo_qsort(m, 0, m->nnz - 1);
  if (merge) {
    int64_t k = 0;
    for (int64_t i = 1; i < m->nnz; i++) {
      if (m->row[k] != m->row[i] || m->col[k] != m->col[i]) {
        k++;
        m->row[k] = m->row[i];
        m->col[k] = m->col[i];
      }
      m->val[k] = m->val[i];
    }
    m->nnz = k + 1;
  }
}

/* util::laplacian_matrix_2D_5p -- src/utils/create_matrix/laplacian_2D5p.cpp:
 * 6-50.  Row ii = i*m + j of the (m*n)^2 matrix; entries inserted in the
 * order (ii-m), (ii-1), ii, (ii+1), (ii+m) when inside the grid. */
void or_laplacian_2d5p(or_coo *out, int m, int n) {
  int64_t size = (int64_t)m * n;
  or_coo_init(out, size, size);
  for (int64_t ii = 0; ii < size; ii++) {
    int64_t i = ii / m;
    int64_t j = ii - i * m;
    if (i > 0)
      or_coo_insert(out, ii, ii - m, -1.0);
    if (j > 0)
      or_coo_insert(out, ii, ii - 1, -1.0);
    or_coo_insert(out, ii, ii, 4.0);
    if (j < m - 1)
      or_coo_insert(out, ii, ii + 1, -1.0);
    if (i < n - 1)
      or_coo_insert(out, ii, ii + m, -1.0);
  }
}

/* COO::input_mm -- src/utils/IO/IO_coo.cpp:69-153.  Banner must read
 * "%%MatrixMarket matrix coordinate real general"; comment lines start with
 * '%'; indices are 1-based. */
int or_input_mm(or_coo *out, const char *path, int is_float) {
  FILE *f = fopen(path, "r");
  if (!f)
    return -1;
  char line[4096];
  char t0[256] = "", t1[256] = "", t2[256] = "", t3[256] = "", t4[256] = "";
  if (!fgets(line, sizeof line, f)) {
    fclose(f);
    return -2;
  }
  sscanf(line, "%255s %255s %255s %255s %255s", t0, t1, t2, t3, t4);
  if (strcmp(t0, "%%MatrixMarket") || strcmp(t1, "matrix") ||
      strcmp(t2, "coordinate") || strcmp(t3, "real") || strcmp(t4, "general")) {
    fclose(f);
    return -3;
  }
  do {
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return -4;
    }
  } while (line[0] == '%');
  long long rows = 0, cols = 0, nnz = 0;
  if (sscanf(line, "%lld %lld %lld", &rows, &cols, &nnz) != 3 || cols <= 0 ||
      nnz < 0) {
    fclose(f);
    return -5;
  }
  or_coo_init(out, rows, cols);
  out->alloc = nnz > 0 ? nnz : 1;
  out->row = (int32_t *)calloc((size_t)out->alloc, sizeof(int32_t));
  out->col = (int32_t *)calloc((size_t)out->alloc, sizeof(int32_t));
  out->val = (double *)calloc((size_t)out->alloc, sizeof(double));
  for (long long k = 0; k < nnz; k++) {
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return -6;
    }
    char *end = NULL;
    long long ix = strtoll(line, &end, 10);
    long long jx = strtoll(end, &end, 10);
    double v = is_float ? (double)strtof(end, NULL) : strtod(end, NULL);
    out->row[k] = (int32_t)(ix - 1);
    out->col[k] = (int32_t)(jx - 1);
    out->val[k] = v;
  }
  out->nnz = nnz;
  fclose(f);
  return 0;
}

/* CRS::convert(COO&) -- src/utils/convert/convert_crs.cpp:9-43.  Values and
 * column indices are copied in COO order.  row_ptr starts all zero; each
 * entry sets row_ptr[row+1] = i+1, rows skipped between two entries get the
 * current offset, and rows after the last non-empty row keep 0 (as the
 * reference leaves them). */
void or_crs_from_coo(const or_coo *coo, int32_t *row_ptr, int32_t *col_ind,
                     double *val) {
  for (int64_t i = 0; i < coo->nnz; i++) {
    val[i] = coo->val[i];
    col_ind[i] = coo->col[i];
  }
  for (int64_t i = 0; i <= coo->rows; i++)
    row_ptr[i] = 0;
  int64_t c_row = 0;
  for (int64_t i = 0; i < coo->nnz; i++) {
    if (c_row == coo->row[i]) {
      row_ptr[c_row + 1] = (int32_t)(i + 1);
    } else {
      for (int64_t j = c_row + 1; j < coo->row[i]; j++)
        row_ptr[j + 1] = (int32_t)i;
      c_row = coo->row[i];
      row_ptr[c_row + 1] = (int32_t)(i + 1);
    }
  }
}

/* ---------------------------------------------------------------------------
 * type-generic kernels and solvers
 * ------------------------------------------------------------------------- */
#define T double
#define S d
#define OR_SQRT sqrt
#include "oracle_t.inc"
#undef T
#undef S
#undef OR_SQRT

#define T float
#define S s
#define OR_SQRT sqrtf
#include "oracle_t.inc"
#undef T
#undef S
#undef OR_SQRT

/* helper for callers that hold COO triplets in their own arrays: sorts (and
 * merges) them in place with or_coo_sort; returns the resulting nnz */
int64_t or_coo_sort_arrays(int64_t nnz, int32_t *row, int32_t *col,
                           double *val, int merge) {
  or_coo m;
  or_coo_init(&m, 0, 0);
  m.nnz = nnz;
  m.alloc = nnz;
  m.row = row;
  m.col = col;
  m.val = val;
  or_coo_sort(&m, merge);
  return m.nnz;
}
