/*
 * CPU ORACLE of the sparse 3D FFT path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or as the reported CPU
 * baseline). The product library (spfft_amd/lib/libspfft.so) never calls it.
 *
 * A plain-C (+ OpenMP) restatement of the reference's SPFFT_PU_HOST pipeline
 * for a local grid, stage by stage:
 *   index map      src/compression/indices.hpp:49-55, :120-186
 *                  src/parameters/parameters.cpp:143-180 (zero-zero stick)
 *   scatter/gather src/compression/compression_host.hpp:51-92
 *   stick symmetry src/symmetry/symmetry_host.hpp:65-90
 *   z transform    src/fft/transform_1d_host.hpp:50-148 (batched per stick)
 *   transpose      src/transpose/transpose_host.hpp:76-155, planes (z, xFreq, y)
 *   plane symmetry src/symmetry/symmetry_host.hpp:43-60
 *   y transform    src/fft/transform_1d_host.hpp:155-231, only distinct x
 *   x transform    C2C src/fft/transform_1d_host.hpp:50-148,
 *                  C2R/R2C src/fft/transform_real_1d_host.hpp:52-241
 *   orchestration  src/execution/execution_host.cpp:50-120, :247-352
 *   scaling        (T)(1.0 / (double)(Nx*Ny*Nz)), execution_host.cpp:54-55
 *
 * The reference delegates the 1D transforms to FFTW (unpinned 3.x; CI uses
 * Ubuntu 22.04's 3.3.8). FFTW computes the unnormalised DFT
 *   X[k] = sum_n x[n] exp(sign 2 pi i k n / N), sign = -1 forward, +1 backward,
 * and its c2r ignores the imaginary parts of the DC and (even N) Nyquist
 * inputs. This file restates that DFT with a recursive mixed-radix
 * Cooley-Tukey (decimation in time, any prime factor by direct summation)
 * and twiddles evaluated in long double, so its rounding is a few ulp of the
 * exact DFT rather than FFTW's bits; tests pin it against numpy's pocketfft,
 * brute-force DFTs and closed forms.
 *
 * Repeated triplets: the reference's OpenMP scatter makes the surviving
 * value unspecified; this oracle keeps the last one (the serial order).
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

enum { ORACLE_OK = 0, ORACLE_INVALID_PARAMETER = 5, ORACLE_INVALID_INDICES = 7 };

/* ------------------------------------------------------------------ FFT */

#define MAX_FACTORS 40

typedef struct {
  int n;
  int nfac;
  int fac[MAX_FACTORS];
  cplx* w; /* w[t] = exp(-2 pi i t / n) */
} fft_plan;

static void fft_plan_init(fft_plan* p, int n) {
  p->n = n;
  p->nfac = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    if ((long long)f * f > m) f = m;
    while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
  }
  p->w = (cplx*)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < n; ++t) {
    /* reduce to (-pi, pi] before evaluating */
    const long long tt = (2LL * t > n) ? (long long)t - n : t;
    const long double a = -twopi * (long double)tt / (long double)n;
    p->w[t].re = (double)cosl(a);
    p->w[t].im = (double)sinl(a);
  }
}

static void fft_plan_free(fft_plan* p) {
  free(p->w);
  p->w = NULL;
}

static inline cplx tw(const fft_plan* p, long long e, int sign) {
  cplx v = p->w[e];
  if (sign > 0) v.im = -v.im;
  return v;
}

/* Recursive decimation in time: out[0..n) = DFT_n of in[0], in[s], in[2s], ...
 * `step` = N / n maps exponents of W_n onto the length-N table. */
static void ct_rec(const cplx* in, size_t s, cplx* out, int n, const int* fac, const fft_plan* p,
                   long long step, int sign, cplx* tmp) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int r = fac[0];
  const int m = n / r;
  for (int q = 0; q < r; ++q) ct_rec(in + q * s, s * r, out + q * m, m, fac + 1, p, step * r, sign, tmp);
  /* X[k + j m] = sum_q W_n^(q (k + j m)) Y_q[k] */
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < r; ++q) tmp[q] = out[q * m + k];
    for (int j = 0; j < r; ++j) {
      const long long kk = k + (long long)j * m;
      cplx acc = tmp[0];
      for (int q = 1; q < r; ++q) {
        const long long e = ((q * kk) % n) * step;
        const cplx w = tw(p, e, sign);
        acc.re += tmp[q].re * w.re - tmp[q].im * w.im;
        acc.im += tmp[q].re * w.im + tmp[q].im * w.re;
      }
      out[kk] = acc;
    }
  }
}

/* out = DFT of the strided line in (out contiguous). tmp >= max factor. */
static void fft_line(const fft_plan* p, const cplx* in, size_t stride, cplx* out, int sign,
                     cplx* tmp) {
  ct_rec(in, stride, out, p->n, p->fac, p, 1, sign, tmp);
}

/* ------------------------------------------- FFTW3 C API, loaded at run time
 *
 * The reference's host path hands its 1D lines to FFTW 3
 * (src/fft/fftw_interface.hpp:41-101, src/fft/fftw_plan_1d.hpp:86-139:
 * fftw_plan_many_dft with FFTW_ESTIMATE, fftw_execute_dft on new arrays). For
 * the CPU baseline timing the oracle can do the same through any library that
 * exports that API (libfftw3.so.3, or MKL's libmkl_rt.so), opened with dlopen
 * by oracle_use_fft_library(). It is off unless enabled: the parity checks use
 * the built-in FFT above, and a CPU test compares the two providers. */
typedef void* (*fftw_plan_many_fn)(int, const int*, int, cplx*, const int*, int, int, cplx*,
                                   const int*, int, int, int, unsigned);
typedef void (*fftw_execute_dft_fn)(void*, cplx*, cplx*);
typedef void (*fftw_destroy_plan_fn)(void*);

static struct {
  void* handle;
  fftw_plan_many_fn plan_many;
  fftw_execute_dft_fn execute_dft;
  fftw_destroy_plan_fn destroy_plan;
} g_fftw;

enum { FFTW_FORWARD_SIGN = -1, FFTW_BACKWARD_SIGN = 1 };
#define FFTW_FLAG_UNALIGNED (1U << 1)
#define FFTW_FLAG_ESTIMATE (1U << 6)

/* 0 on success; later plans use the library. NULL turns it off again. */
int oracle_use_fft_library(const char* path) {
  if (!path) {
    g_fftw.plan_many = NULL;
    return 0;
  }
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 1;
  fftw_plan_many_fn pm = (fftw_plan_many_fn)dlsym(h, "fftw_plan_many_dft");
  fftw_execute_dft_fn ex = (fftw_execute_dft_fn)dlsym(h, "fftw_execute_dft");
  fftw_destroy_plan_fn de = (fftw_destroy_plan_fn)dlsym(h, "fftw_destroy_plan");
  if (!pm || !ex || !de) {
    dlclose(h);
    return 2;
  }
  g_fftw.handle = h;
  g_fftw.plan_many = pm;
  g_fftw.execute_dft = ex;
  g_fftw.destroy_plan = de;
  return 0;
}

/* One out-of-place line plan: length n, input stride, unit output stride. */
static void* ext_plan(int n, int istride, int sign, cplx* in, cplx* out) {
  if (!g_fftw.plan_many) return NULL;
  const int dims[1] = {n};
  return g_fftw.plan_many(1, dims, 1, in, NULL, istride, n * istride, out, NULL, 1, n, sign,
                          FFTW_FLAG_ESTIMATE | FFTW_FLAG_UNALIGNED);
}

/* ------------------------------------------------------------ index map */

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

static int storage(int dim, int i) { return i < 0 ? dim + i : i; }

/* Returns ORACLE_OK and fills valueSlot[n], stickKey[<=n], *numSticks. */
int oracle_index_map(int r2c, int nx, int ny, int nz, int n, const int* t, int* valueSlot,
                     int* stickKey, int* numSticks) {
  if ((unsigned long long)n > (unsigned long long)nx * (unsigned long long)ny * (unsigned long long)nz)
    return ORACLE_INVALID_PARAMETER;
  int centered = 0;
  for (long long i = 0; i < 3LL * n; ++i)
    if (t[i] < 0) { centered = 1; break; }
  const int maxX = (r2c || centered ? nx / 2 + 1 : nx) - 1;
  const int maxY = (centered ? ny / 2 + 1 : ny) - 1;
  const int maxZ = (centered ? nz / 2 + 1 : nz) - 1;
  const int minX = r2c ? 0 : maxX - nx + 1;
  const int minY = maxY - ny + 1;
  const int minZ = maxZ - nz + 1;
  for (int i = 0; i < n; ++i) {
    const int x = t[3 * i], y = t[3 * i + 1], z = t[3 * i + 2];
    if (x < minX || x > maxX || y < minY || y > maxY || z < minZ || z > maxZ)
      return ORACLE_INVALID_INDICES;
  }
  int* keys = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) keys[i] = storage(nx, t[3 * i]) * ny + storage(ny, t[3 * i + 1]);
  memcpy(stickKey, keys, sizeof(int) * (size_t)n);
  qsort(stickKey, (size_t)n, sizeof(int), cmp_int);
  int ns = 0;
  for (int i = 0; i < n; ++i)
    if (ns == 0 || stickKey[ns - 1] != stickKey[i]) stickKey[ns++] = stickKey[i];
  for (int i = 0; i < n; ++i) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (stickKey[mid] < keys[i]) lo = mid + 1; else hi = mid;
    }
    valueSlot[i] = lo * nz + storage(nz, t[3 * i + 2]);
  }
  *numSticks = ns;
  free(keys);
  return ORACLE_OK;
}

/* ----------------------------------------------------------------- plan */

typedef struct {
  int r2c, nx, ny, nz, nxf, nv, nst, nuniq, zeroZeroStick, nthreads;
  int* valueSlot;
  int* stickKey;
  int* uniqX;
  fft_plan px, py, pz;
  cplx* sticks; /* [stick][z] */
  cplx* planes; /* [z][xFreq][y] */
  int maxlen;
  void* ext[7]; /* FFTW3-API line plans (EXT_*), all NULL with the built-in FFT */
} oracle_plan;

enum { EXT_Z_BWD, EXT_Y_BWD, EXT_Y_FWD, EXT_X_BWD_STRIDED, EXT_X_BWD, EXT_X_FWD, EXT_Z_FWD };

/* out = DFT of the strided line `in`, by the plan's FFTW3-API line plan
 * `which` when there is one, else by the built-in FFT. */
static void run_line(const oracle_plan* p, int which, const fft_plan* fp, const cplx* in,
                     size_t stride, cplx* out, int sign, cplx* tmp) {
  if (p->ext[which]) {
    g_fftw.execute_dft(p->ext[which], (cplx*)in, out);
  } else {
    fft_line(fp, in, stride, out, sign, tmp);
  }
}

void oracle_plan_destroy(oracle_plan* p) {
  if (!p) return;
  free(p->valueSlot);
  free(p->stickKey);
  free(p->uniqX);
  free(p->sticks);
  free(p->planes);
  fft_plan_free(&p->px);
  fft_plan_free(&p->py);
  fft_plan_free(&p->pz);
  for (int k = 0; k < 7; ++k)
    if (p->ext[k]) g_fftw.destroy_plan(p->ext[k]);
  free(p);
}

/* Returns NULL and sets *err on invalid input. nthreads < 1: OpenMP default. */
oracle_plan* oracle_plan_create(int r2c, int nx, int ny, int nz, int n, const int* triplets,
                                int nthreads, int* err) {
  *err = ORACLE_OK;
  if (nx <= 0 || ny <= 0 || nz <= 0 || n < 0 || (!triplets && n > 0)) {
    *err = ORACLE_INVALID_PARAMETER;
    return NULL;
  }
  oracle_plan* p = (oracle_plan*)calloc(1, sizeof(oracle_plan));
  p->r2c = r2c;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->nxf = r2c ? nx / 2 + 1 : nx;
  p->nv = n;
  p->nthreads = nthreads;
  p->valueSlot = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  p->stickKey = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  *err = oracle_index_map(r2c, nx, ny, nz, n, triplets, p->valueSlot, p->stickKey, &p->nst);
  if (*err != ORACLE_OK) {
    free(p->valueSlot);
    free(p->stickKey);
    free(p);
    return NULL;
  }
  p->zeroZeroStick = (p->nst > 0 && p->stickKey[0] == 0) ? 0 : p->nst;
  /* distinct x of the sticks, ascending (execution_host.cpp:60-63) */
  p->uniqX = (int*)malloc(sizeof(int) * (size_t)(p->nst > 0 ? p->nst : 1));
  p->nuniq = 0;
  for (int s = 0; s < p->nst; ++s) {
    const int x = p->stickKey[s] / ny;
    if (p->nuniq == 0 || p->uniqX[p->nuniq - 1] != x) p->uniqX[p->nuniq++] = x;
  }
  fft_plan_init(&p->px, nx);
  fft_plan_init(&p->py, ny);
  fft_plan_init(&p->pz, nz);
  p->sticks = (cplx*)malloc(sizeof(cplx) * (size_t)(p->nst > 0 ? p->nst : 1) * nz);
  p->planes = (cplx*)malloc(sizeof(cplx) * (size_t)nz * p->nxf * ny);
  p->maxlen = nx > ny ? nx : ny;
  if (nz > p->maxlen) p->maxlen = nz;
  if (g_fftw.plan_many) {
    /* planning only (FFTW_ESTIMATE reads no data); executed on new arrays */
    cplx* a = (cplx*)calloc((size_t)nz * p->nxf * ny + 1, sizeof(cplx));
    cplx* b = (cplx*)calloc((size_t)p->maxlen + 1, sizeof(cplx));
    p->ext[EXT_Z_BWD] = ext_plan(nz, 1, FFTW_BACKWARD_SIGN, a, b);
    p->ext[EXT_Y_BWD] = ext_plan(ny, 1, FFTW_BACKWARD_SIGN, a, b);
    p->ext[EXT_Y_FWD] = ext_plan(ny, 1, FFTW_FORWARD_SIGN, a, b);
    p->ext[EXT_X_BWD_STRIDED] = ext_plan(nx, ny, FFTW_BACKWARD_SIGN, a, b);
    p->ext[EXT_X_BWD] = ext_plan(nx, 1, FFTW_BACKWARD_SIGN, a, b);
    p->ext[EXT_X_FWD] = ext_plan(nx, 1, FFTW_FORWARD_SIGN, a, b);
    p->ext[EXT_Z_FWD] = ext_plan(nz, p->nxf * ny, FFTW_FORWARD_SIGN, a, b);
    free(a);
    free(b);
  }
  return p;
}

int oracle_plan_num_sticks(const oracle_plan* p) { return p->nst; }

#ifdef _OPENMP
#define ORACLE_PARALLEL _Pragma("omp parallel num_threads(nthr)")
#define ORACLE_FOR _Pragma("omp for schedule(static)")
#else
#define ORACLE_PARALLEL
#define ORACLE_FOR
#endif

static int thread_count(const oracle_plan* p) {
#ifdef _OPENMP
  return p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
  (void)p;
  return 1;
#endif
}

static inline int nonzero(cplx v) { return v.re != 0.0 || v.im != 0.0; }

/* Hermitian fill of one line with the reference's serial rule
 * (symmetry_host.hpp:47-58 / :72-89): walk k = 1..n-1 in order, copy the
 * conjugate of every non-zero entry to n-k. */
static void hermitian_fill(cplx* v, int n, size_t stride) {
  for (int k = 1; k < n / 2 + 1; ++k) {
    const cplx a = v[k * stride];
    if (nonzero(a)) { v[(n - k) * stride].re = a.re; v[(n - k) * stride].im = -a.im; }
  }
  for (int k = n / 2 + 1; k < n; ++k) {
    const cplx a = v[k * stride];
    if (nonzero(a)) { v[(n - k) * stride].re = a.re; v[(n - k) * stride].im = -a.im; }
  }
}

/* values: 2*nv doubles (re, im). out: space domain, (z*ny + y)*nx + x,
 * complex interleaved (C2C) or real (R2C). */
void oracle_backward(oracle_plan* p, const double* values, double* out) {
  const int nx = p->nx, ny = p->ny, nz = p->nz, nxf = p->nxf, nst = p->nst;
  const int nthr = thread_count(p);
  const cplx* in = (const cplx*)values;
  /* scatter into zeroed sticks */
  memset(p->sticks, 0, sizeof(cplx) * (size_t)nst * nz);
  for (int i = 0; i < p->nv; ++i) p->sticks[p->valueSlot[i]] = in[i];
  if (p->r2c && p->zeroZeroStick < nst) hermitian_fill(p->sticks + (size_t)p->zeroZeroStick * nz, nz, 1);
  memset(p->planes, 0, sizeof(cplx) * (size_t)nz * nxf * ny);
  ORACLE_PARALLEL {
    cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)p->maxlen);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)p->maxlen);
    /* z transform, then transpose into planes[z][x][y] */
    ORACLE_FOR
    for (int s = 0; s < nst; ++s) {
      run_line(p, EXT_Z_BWD, &p->pz, p->sticks + (size_t)s * nz, 1, line, +1, tmp);
      const int x = p->stickKey[s] / ny, y = p->stickKey[s] % ny;
      for (int z = 0; z < nz; ++z) p->planes[((size_t)z * nxf + x) * ny + y] = line[z];
    }
    /* plane symmetry on x = 0 (applied to every z plane) */
    if (p->r2c) {
      ORACLE_FOR
      for (int z = 0; z < nz; ++z) hermitian_fill(p->planes + (size_t)z * nxf * ny, ny, 1);
    }
    /* y transform on distinct x only */
    ORACLE_FOR
    for (long long zu = 0; zu < (long long)nz * p->nuniq; ++zu) {
      const int z = (int)(zu / p->nuniq), x = p->uniqX[zu % p->nuniq];
      cplx* col = p->planes + ((size_t)z * nxf + x) * ny;
      run_line(p, EXT_Y_BWD, &p->py, col, 1, line, +1, tmp);
      memcpy(col, line, sizeof(cplx) * (size_t)ny);
    }
    /* x transform of every row into the space domain */
    cplx* full = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
    ORACLE_FOR
    for (long long zy = 0; zy < (long long)nz * ny; ++zy) {
      const int z = (int)(zy / ny), y = (int)(zy % ny);
      const cplx* src = p->planes + (size_t)z * nxf * ny + y;
      if (!p->r2c) {
        run_line(p, EXT_X_BWD_STRIDED, &p->px, src, (size_t)ny, (cplx*)out + (size_t)zy * nx, +1, tmp);
      } else {
        for (int k = 0; k < nxf; ++k) full[k] = src[(size_t)k * ny];
        full[0].im = 0.0;
        if (nx % 2 == 0) full[nx / 2].im = 0.0;
        for (int k = nxf; k < nx; ++k) { full[k].re = full[nx - k].re; full[k].im = -full[nx - k].im; }
        run_line(p, EXT_X_BWD, &p->px, full, 1, line, +1, tmp);
        for (int x = 0; x < nx; ++x) out[(size_t)zy * nx + x] = line[x].re;
      }
    }
    free(full);
    free(line);
    free(tmp);
  }
}

/* space: space domain as above. out: 2*nv doubles. scaling != 0 applies
 * (double)(1.0 / (Nx*Ny*Nz)) to each component. */
void oracle_forward(oracle_plan* p, const double* space, double* out, int scaling) {
  const int nx = p->nx, ny = p->ny, nz = p->nz, nxf = p->nxf, nst = p->nst;
  const int nthr = thread_count(p);
  ORACLE_PARALLEL {
    cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)p->maxlen);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)p->maxlen);
    cplx* real = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
    /* x transform of every row into planes[z][x][y] */
    ORACLE_FOR
    for (long long zy = 0; zy < (long long)nz * ny; ++zy) {
      const int z = (int)(zy / ny), y = (int)(zy % ny);
      if (!p->r2c) {
        run_line(p, EXT_X_FWD, &p->px, (const cplx*)space + (size_t)zy * nx, 1, line, -1, tmp);
      } else {
        for (int x = 0; x < nx; ++x) { real[x].re = space[(size_t)zy * nx + x]; real[x].im = 0.0; }
        run_line(p, EXT_X_FWD, &p->px, real, 1, line, -1, tmp);
      }
      cplx* dst = p->planes + (size_t)z * nxf * ny + y;
      for (int k = 0; k < nxf; ++k) dst[(size_t)k * ny] = line[k];
    }
    /* y transform on distinct x only */
    ORACLE_FOR
    for (long long zu = 0; zu < (long long)nz * p->nuniq; ++zu) {
      const int z = (int)(zu / p->nuniq), x = p->uniqX[zu % p->nuniq];
      cplx* col = p->planes + ((size_t)z * nxf + x) * ny;
      run_line(p, EXT_Y_FWD, &p->py, col, 1, line, -1, tmp);
      memcpy(col, line, sizeof(cplx) * (size_t)ny);
    }
    /* gather sticks and z transform */
    ORACLE_FOR
    for (int s = 0; s < nst; ++s) {
      const int x = p->stickKey[s] / ny, y = p->stickKey[s] % ny;
      run_line(p, EXT_Z_FWD, &p->pz, p->planes + (size_t)x * ny + y, (size_t)nxf * ny,
               p->sticks + (size_t)s * nz, -1, tmp);
    }
    free(real);
    free(line);
    free(tmp);
  }
  const double scale = 1.0 / (double)((unsigned long long)nx * (unsigned long long)ny * (unsigned long long)nz);
  cplx* o = (cplx*)out;
  for (int i = 0; i < p->nv; ++i) {
    cplx v = p->sticks[p->valueSlot[i]];
    if (scaling) { v.re *= scale; v.im *= scale; }
    o[i] = v;
  }
}

/* Plain 1D DFT through the same FFT, exported so tests can pin it against
 * numpy and brute force. sign: -1 forward, +1 backward. */
void oracle_fft1d(int n, const double* in, double* out, int sign) {
  fft_plan p;
  fft_plan_init(&p, n);
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  fft_line(&p, (const cplx*)in, 1, (cplx*)out, sign, tmp);
  free(tmp);
  fft_plan_free(&p);
}
