/* sanitize_check.c -- runs the oracle's C loops under AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`).  Test
 * infrastructure only, like the oracle itself.
 *
 * The inputs are built the way the oracle's Python side builds them
 * (meshes.py:70-186 conventions): padded (ny+6) x (nx+6) grids, closed and
 * periodic directions, masks with random circular islands, the derived
 * masks and stencil orders, random fields.  Every flux method runs through
 * compflux, vortexforce and innerproduct over the whole flat range, as
 * oracle.py calls them; then the index map and the Laplacian assembly.  The
 * flat-index loops read neighbours without an upper bound check (numba's
 * semantics): the stencil orders are what keeps them in bounds, and this
 * program is where that is checked.
 *
 * Built with -DPLANT_OOB the program plants one order one step too high at
 * the end of the array: the sanitizer must then report a heap overflow
 * (the check of the check). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

double orc_weno3z(double qm, double q0, double qp);
double orc_weno5z(double qmm, double qm, double q0, double qp, double qpp);
double orc_f1(int method, double U, double qm, double qp);
double orc_f3(int method, double U, double qmm, double qm, double qp, double qpp);
double orc_f5(int method, double U, double qmmm, double qmm, double qm, double qp,
              double qpp, double qppp);
void orc_compflux(double* flx, const double* U, const double* q, const int8_t* o,
                  int64_t n, int64_t s, int method, int64_t i0, int64_t i1);
void orc_vortexforce(double* du, const double* V, const double* q, const int8_t* o,
                     int64_t n, int64_t s, int64_t s2, int sign, int method, int64_t i0,
                     int64_t i1);
void orc_innerproduct(double* ke, const double* U, const double* q, const int8_t* o,
                      int64_t n, int64_t s, int method, int64_t i0, int64_t i1);
void orc_set_order(const int8_t* m, int64_t n, int64_t shift, int maxorder, int8_t* o);
void orc_set_masks(const int8_t* msk, int64_t ny, int64_t nx, int8_t* mskx, int8_t* msky,
                   int8_t* mskv);
int64_t orc_gindex(const int8_t* msk, int64_t n, int32_t* G);
int64_t orc_laplacian_coo(const int32_t* G, int64_t ny, int64_t nx, double dx, double dy,
                          double maindiag, int neumann, int xperiodic, int yperiodic,
                          int halowidth, int32_t* row, int32_t* col, double* data);

static uint64_t rs = 0x9E3779B97F4A7C15ull;
static double urand(void) { /* xorshift64*, in [0, 1) */
  rs ^= rs >> 12;
  rs ^= rs << 25;
  rs ^= rs >> 27;
  return (double)((rs * 0x2545F4914F6CDD1Dull) >> 11) / 9007199254740992.0;
}

static void* xalloc(size_t n) {
  void* p = calloc(n, 1);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

static double checksum(const double* a, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += a[i];
  return s;
}

/* one grid: returns 0 when every output is finite */
static int run_case(int ny, int nx, int xper, int yper, int nisl) {
  const int h = 3, NY = ny + 2 * h, NX = nx + 2 * h;
  const int64_t n = (int64_t)NY * NX;
  int8_t* msk = xalloc(n);
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      const int inx = xper || (i >= h && i < NX - h);
      const int iny = yper || (j >= h && j < NY - h);
      msk[(int64_t)j * NX + i] = (int8_t)(inx && iny);
    }
  for (int k = 0; k < nisl; k++) { /* circular islands */
    const double cj = h + urand() * ny, ci = h + urand() * nx, r = 2 + urand() * ny / 6.0;
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++)
        if ((j - cj) * (j - cj) + (i - ci) * (i - ci) < r * r) msk[(int64_t)j * NX + i] = 0;
  }
  if (xper) /* periodic images of the halo columns */
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < h; i++) {
        msk[(int64_t)j * NX + i] = msk[(int64_t)j * NX + i + nx];
        msk[(int64_t)j * NX + NX - h + i] = msk[(int64_t)j * NX + h + i];
      }
  if (yper)
    for (int j = 0; j < h; j++)
      for (int i = 0; i < NX; i++) {
        msk[(int64_t)j * NX + i] = msk[(int64_t)(j + ny) * NX + i];
        msk[(int64_t)(NY - h + j) * NX + i] = msk[(int64_t)(h + j) * NX + i];
      }
  int8_t *mx = xalloc(n), *my = xalloc(n), *mv = xalloc(n);
  orc_set_masks(msk, NY, NX, mx, my, mv);
  int8_t *ocx = xalloc(n), *ocy = xalloc(n), *ovx = xalloc(n), *ovy = xalloc(n);
  int8_t *okx = xalloc(n), *oky = xalloc(n);
  orc_set_order(msk, n, 1, 6, ocx);
  orc_set_order(msk, n, NX, 6, ocy);
  orc_set_order(mv, n, -1, 6, ovx);
  orc_set_order(mv, n, -NX, 6, ovy);
  for (int64_t k = 0; k < n; k++) {
    ovx[k] = (int8_t)(ovx[k] * my[k]);
    ovy[k] = (int8_t)(ovy[k] * mx[k]);
  }
  orc_set_order(mx, n, -1, 6, okx);
  orc_set_order(my, n, -NX, 6, oky);
#ifdef PLANT_OOB
  ocy[n - 1] = 6; /* a sixth-order stencil reading two rows past the end */
#endif
  double *q = xalloc(8 * n), *U = xalloc(8 * n), *V = xalloc(8 * n), *out = xalloc(8 * n);
  for (int64_t k = 0; k < n; k++) {
    q[k] = urand() - 0.5;
    U[k] = 2 * urand() - 1;
    V[k] = 2 * urand() - 1;
  }
  int bad = 0;
  for (int method = 0; method < 4; method++) {
    orc_compflux(out, U, q, ocx, n, 1, method, 0, n);
    bad |= !isfinite(checksum(out, n));
    orc_compflux(out, U, q, ocy, n, NX, method, 0, n);
    bad |= !isfinite(checksum(out, n));
    orc_vortexforce(out, V, q, ovy, n, NX, 1, +1, method, 0, n);
    bad |= !isfinite(checksum(out, n));
    orc_vortexforce(out, V, q, ovx, n, 1, NX, -1, method, 0, n);
    bad |= !isfinite(checksum(out, n));
    memset(out, 0, 8 * n);
    orc_innerproduct(out, U, q, okx, n, 1, method, 0, n);
    orc_innerproduct(out, U, q, oky, n, NX, method, 0, n);
    bad |= !isfinite(checksum(out, n));
  }
  /* the Poisson operator of the vertex unknowns (halo-stripped mskv) */
  int8_t* inner = xalloc(n);
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      const int keepx = xper ? (i >= h && i < NX - h) : 1;
      const int keepy = yper ? (j >= h && j < NY - h) : 1;
      inner[(int64_t)j * NX + i] = (int8_t)(mv[(int64_t)j * NX + i] && keepx && keepy);
    }
  int32_t* G = xalloc(4 * n);
  const int64_t nu = orc_gindex(inner, n, G);
  int32_t *row = xalloc(4 * 5 * (size_t)(nu + 1)), *col = xalloc(4 * 5 * (size_t)(nu + 1));
  double* data = xalloc(8 * 5 * (size_t)(nu + 1));
  for (int neu = 0; neu < 2; neu++) {
    const int64_t nt = orc_laplacian_coo(G, NY, NX, 1.0 / nx, 1.3 / ny, 0.2, neu, xper, yper,
                                         h, row, col, data);
    bad |= nt > 5 * nu;
    for (int64_t k = 0; k < nt; k++) bad |= row[k] < 0 || row[k] >= nu || col[k] < 0 || col[k] >= nu;
  }
  free(msk), free(mx), free(my), free(mv), free(ocx), free(ocy), free(ovx), free(ovy);
  free(okx), free(oky), free(q), free(U), free(V), free(out), free(inner), free(G);
  free(row), free(col), free(data);
  return bad;
}

int main(void) {
  int bad = 0, cases = 0;
  /* the reconstructions on their own, flat and extreme inputs included */
  const double vals[] = {0.0, 1.0, -1.0, 1e-300, 1e30, 0.5, -0.25};
  for (int a = 0; a < 7; a++)
    for (int b = 0; b < 7; b++)
      for (int c = 0; c < 7; c++) {
        bad |= isnan(orc_weno3z(vals[a], vals[b], vals[c]));
        bad |= isnan(orc_weno5z(vals[a], vals[b], vals[c], vals[b], vals[a]));
        for (int m = 0; m < 4; m++) {
          bad |= isnan(orc_f1(m, vals[a] - 0.5, vals[b], vals[c]));
          bad |= isnan(orc_f3(m, vals[c] - 0.5, vals[a], vals[b], vals[c], vals[a]));
          bad |= isnan(orc_f5(m, vals[b] - 0.5, vals[a], vals[b], vals[c], vals[a], vals[b],
                              vals[c]));
        }
      }
  const int grids[][5] = {{16, 16, 0, 0, 0}, {24, 40, 1, 0, 3}, {40, 24, 0, 1, 3},
                          {32, 32, 1, 1, 4}, {37, 70, 0, 0, 5}, {64, 48, 1, 0, 6},
                          {12, 9, 1, 1, 1}};
  for (size_t g = 0; g < sizeof(grids) / sizeof(grids[0]); g++) {
    bad |= run_case(grids[g][0], grids[g][1], grids[g][2], grids[g][3], grids[g][4]);
    cases++;
  }
  if (bad) {
    printf("sanitize check: non-finite or out-of-range output\n");
    return 1;
  }
  printf("sanitize check: %d grids clean\n", cases);
  return 0;
}
