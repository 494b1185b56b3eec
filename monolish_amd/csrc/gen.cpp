// gen.cpp -- host generators of the benchmark workloads (3D stencils, random
// CRS with lognormal row lengths).  Deterministic for a given seed and
// independent of the number of host threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../include/monolish_amd.h"

namespace {

template <class F> void par_for(int64_t n, F fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<int64_t>(std::max(1u, std::min(hw, 16u)), std::max<int64_t>(1, n / 4096));
  if (nt <= 1) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back(fn, b, e);
  }
  for (auto &t : th) t.join();
}

inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in [0, 1) with 53 random bits
inline double u01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

struct Stencil {
  int64_t nx, ny, nz;
  int points;
  int count(int64_t i, int64_t j, int64_t k) const {
    if (points == 7) {
      return 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
    }
    const int ci = 1 + (i > 0) + (i < nx - 1), cj = 1 + (j > 0) + (j < ny - 1),
              ck = 1 + (k > 0) + (k < nz - 1);
    return ci * cj * ck;
  }
};

}  // namespace

extern "C" {

int mgx_gen_stencil3d(int64_t nx, int64_t ny, int64_t nz, int points, double shift,
                      int64_t *nnz_out, int32_t *row_ptr, int32_t *col_ind, double *val) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || (points != 7 && points != 27) || !nnz_out)
    return MGX_ERR_INVALID;
  const int64_t n = nx * ny * nz;
  int64_t nnz;
  if (points == 7)
    nnz = 7 * n - 2 * (ny * nz + nx * nz + nx * ny);
  else
    nnz = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
  *nnz_out = nnz;
  if (n >= INT32_MAX || nnz >= INT32_MAX) return MGX_ERR_INVALID;
  if (!row_ptr) return MGX_OK;
  const Stencil st{nx, ny, nz, points};
  // row lengths per z-plane, then a prefix over planes
  std::vector<int64_t> plane(nz + 1, 0);
  par_for(nz, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; k++) {
      int64_t s = 0;
      for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) s += st.count(i, j, k);
      plane[k + 1] = s;
    }
  });
  for (int64_t k = 0; k < nz; k++) plane[k + 1] += plane[k];
  const double diag = (points == 7 ? 6.0 : 26.0) + shift;
  par_for(nz, [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; k++) {
      int64_t e = plane[k];
      for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
          const int64_t row = (k * ny + j) * nx + i;
          row_ptr[row] = (int32_t)e;
          for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) {
                const int64_t kk = k + dk, jj = j + dj, ii = i + di;
                if (kk < 0 || kk >= nz || jj < 0 || jj >= ny || ii < 0 || ii >= nx) continue;
                const int nzc = (dk != 0) + (dj != 0) + (di != 0);
                if (points == 7 && nzc > 1) continue;
                if (col_ind) {
                  col_ind[e] = (int32_t)((kk * ny + jj) * nx + ii);
                  val[e] = nzc == 0 ? diag : -1.0;
                }
                e++;
              }
        }
    }
  });
  row_ptr[n] = (int32_t)nnz;
  return MGX_OK;
}

int mgx_gen_random_lognormal(int64_t rows, int64_t cols, double mean, double sigma,
                             int32_t maxlen, uint64_t seed, int32_t *row_ptr,
                             int32_t *col_ind, double *val) {
  if (rows < 0 || cols <= 0 || mean <= 0 || sigma < 0 || maxlen < 1 || !row_ptr)
    return MGX_ERR_INVALID;
  const double mu = std::log(mean) - 0.5 * sigma * sigma;
  const int64_t cap = std::min<int64_t>(maxlen, cols);
  std::vector<int32_t> len(rows);
  par_for(rows, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; r++) {
      uint64_t s = seed * 0x100000001B3ull ^ (uint64_t)r * 0xD1B54A32D192ED03ull;
      const double u1 = 1.0 - u01(s), u2 = u01(s);
      const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      double L = std::round(std::exp(mu + sigma * g));
      L = std::min<double>(std::max(L, 1.0), (double)cap);
      len[r] = (int32_t)L;
    }
  });
  int64_t acc = 0;
  row_ptr[0] = 0;
  for (int64_t r = 0; r < rows; r++) {
    acc += len[r];
    if (acc >= INT32_MAX) return MGX_ERR_INVALID;
    row_ptr[r + 1] = (int32_t)acc;
  }
  if (!col_ind || !val) return MGX_OK;
  par_for(rows, [&](int64_t b, int64_t e) {
    std::vector<int32_t> tmp;
    for (int64_t r = b; r < e; r++) {
      uint64_t s = (seed + 0x632BE59BD9B4E019ull) * 0x9E3779B97F4A7C15ull ^
                   (uint64_t)r * 0xC2B2AE3D27D4EB4Full;
      const int32_t L = len[r];
      tmp.clear();
      while ((int32_t)tmp.size() < L) {
        while ((int32_t)tmp.size() < L)
          tmp.push_back((int32_t)(splitmix64(s) % (uint64_t)cols));
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      }
      const int64_t off = row_ptr[r];
      for (int32_t j = 0; j < L; j++) {
        col_ind[off + j] = tmp[j];
        val[off + j] = 2.0 * u01(s) - 1.0;
      }
    }
  });
  return MGX_OK;
}

// util::laplacian_matrix_2D_5p(m, n) (src/utils/create_matrix/
// laplacian_2D5p.cpp:6-50) converted to CRS: row ii = i*m + j with entries
// (ii-m), (ii-1), ii, (ii+1), (ii+m) inside the grid, already in column order
int mgx_gen_laplacian2d(int64_t m, int64_t n, int64_t *nnz_out, int32_t *row_ptr,
                        int32_t *col_ind, double *val) {
  if (m <= 0 || n <= 0 || !nnz_out) return MGX_ERR_INVALID;
  const int64_t size = m * n, nnz = 5 * m * n - 2 * m - 2 * n;
  *nnz_out = nnz;
  if (size >= INT32_MAX || nnz >= INT32_MAX) return MGX_ERR_INVALID;
  if (!row_ptr) return MGX_OK;
  int64_t e = 0;
  for (int64_t ii = 0; ii < size; ii++) {
    const int64_t i = ii / m, j = ii - i * m;
    row_ptr[ii] = (int32_t)e;
    auto put = [&](int64_t c, double v) {
      if (col_ind) {
        col_ind[e] = (int32_t)c;
        val[e] = v;
      }
      e++;
    };
    if (i > 0) put(ii - m, -1.0);
    if (j > 0) put(ii - 1, -1.0);
    put(ii, 4.0);
    if (j < m - 1) put(ii + 1, -1.0);
    if (i < n - 1) put(ii + m, -1.0);
  }
  row_ptr[size] = (int32_t)e;
  return MGX_OK;
}

// 2D convection-diffusion, -lap(u) + pe * du/dx on an m x m grid, first-order
// upwind (flow in +x): diag 4 + c, west -1 - c, east -1, south/north -1 with
// c = pe / (m + 1); nonsymmetric, diagonally dominant
int mgx_gen_convdiff2d(int64_t m, double peclet, int64_t *nnz_out, int32_t *row_ptr,
                       int32_t *col_ind, double *val) {
  if (m <= 0 || peclet < 0 || !nnz_out) return MGX_ERR_INVALID;
  const int64_t size = m * m, nnz = 5 * m * m - 4 * m;
  *nnz_out = nnz;
  if (size >= INT32_MAX || nnz >= INT32_MAX) return MGX_ERR_INVALID;
  if (!row_ptr) return MGX_OK;
  const double c = peclet / (double)(m + 1);
  int64_t e = 0;
  for (int64_t ii = 0; ii < size; ii++) {
    const int64_t i = ii / m, j = ii - i * m;
    row_ptr[ii] = (int32_t)e;
    auto put = [&](int64_t col, double v) {
      if (col_ind) {
        col_ind[e] = (int32_t)col;
        val[e] = v;
      }
      e++;
    };
    if (i > 0) put(ii - m, -1.0);
    if (j > 0) put(ii - 1, -1.0 - c);
    put(ii, 4.0 + c);
    if (j < m - 1) put(ii + 1, -1.0);
    if (i < m - 1) put(ii + m, -1.0);
  }
  row_ptr[size] = (int32_t)e;
  return MGX_OK;
}

// SPD M-matrix: each row draws k in [kmin, kmax] distinct off-diagonal
// columns with values U[-1, 0); the pattern is symmetrised as (A + A^T)/2 on
// the union of both patterns, and diag = sum |off-diagonal| + shift.  Two
// calls: row_ptr only (col_ind/val NULL) gives nnz, then the fill.
int mgx_gen_mmatrix(int64_t n, int32_t kmin, int32_t kmax, double shift, uint64_t seed,
                    int32_t *row_ptr, int32_t *col_ind, double *val) {
  if (n <= 1 || kmin < 0 || kmax < kmin || !row_ptr) return MGX_ERR_INVALID;
  struct Ent {
    int32_t r, c;
    double v;
  };
  std::vector<Ent> ents;
  ents.reserve((size_t)n * (kmax + 1) * 2);
  std::vector<int32_t> tmp;
  for (int64_t r = 0; r < n; r++) {
    uint64_t s = (seed + 0x51ED27ull) * 0x9E3779B97F4A7C15ull ^ (uint64_t)r * 0xD6E8FEB86659FD93ull;
    const int32_t k = kmin + (int32_t)(splitmix64(s) % (uint64_t)(kmax - kmin + 1));
    const int32_t kk = (int32_t)std::min<int64_t>(k, n - 1);
    tmp.clear();
    while ((int32_t)tmp.size() < kk) {
      int32_t c = (int32_t)(splitmix64(s) % (uint64_t)n);
      if (c == r || std::find(tmp.begin(), tmp.end(), c) != tmp.end()) continue;
      tmp.push_back(c);
    }
    for (int32_t c : tmp) {
      const double v = -u01(s);  // U(-1, 0]
      ents.push_back({(int32_t)r, c, 0.5 * v});
      ents.push_back({c, (int32_t)r, 0.5 * v});
    }
  }
  std::sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  // merge duplicates (summed), add the diagonal
  std::vector<Ent> out;
  out.reserve(ents.size() + n);
  std::vector<double> absum(n, 0.0);
  for (size_t i = 0; i < ents.size();) {
    size_t j = i;
    double v = 0;
    while (j < ents.size() && ents[j].r == ents[i].r && ents[j].c == ents[i].c) v += ents[j++].v;
    out.push_back({ents[i].r, ents[i].c, v});
    absum[ents[i].r] += std::fabs(v);
    i = j;
  }
  std::vector<Ent> full;
  full.reserve(out.size() + n);
  size_t p = 0;
  for (int64_t r = 0; r < n; r++) {
    while (p < out.size() && out[p].r == r && out[p].c < r) full.push_back(out[p++]);
    full.push_back({(int32_t)r, (int32_t)r, absum[r] + shift});
    while (p < out.size() && out[p].r == r) full.push_back(out[p++]);
  }
  if ((int64_t)full.size() >= INT32_MAX) return MGX_ERR_INVALID;
  int64_t e = 0;
  for (int64_t r = 0; r < n; r++) {
    row_ptr[r] = (int32_t)e;
    while (e < (int64_t)full.size() && full[e].r == r) {
      if (col_ind) {
        col_ind[e] = full[e].c;
        val[e] = full[e].v;
      }
      e++;
    }
  }
  row_ptr[n] = (int32_t)e;
  return MGX_OK;
}

// vector<T>(N, min, max, seed) and util::random_vector
// (src/utils/vector_constructor.cpp:78-113): std::mt19937 seeded with seed,
// std::uniform_real_distribution<double>(min, max), drawn in index order.
// The FP32 forms pass float bounds to the double distribution and store
// each double draw rounded to float.
int mgx_gen_random_vector(int64_t n, double min, double max, uint32_t seed, mgx_dtype dtype,
                          void *out) {
  if (n < 0 || (n > 0 && !out) || (dtype != MGX_F64 && dtype != MGX_F32)) return MGX_ERR_INVALID;
  std::mt19937 mt(seed);
  if (dtype == MGX_F64) {
    std::uniform_real_distribution<> rand(min, max);
    double *o = static_cast<double *>(out);
    for (int64_t i = 0; i < n; i++) o[i] = rand(mt);
  } else {
    std::uniform_real_distribution<> rand((float)min, (float)max);
    float *o = static_cast<float *>(out);
    for (int64_t i = 0; i < n; i++) o[i] = (float)rand(mt);
  }
  return MGX_OK;
}

}  // extern "C"
