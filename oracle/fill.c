/* Oracle (test infrastructure only): CPU restatement of the fused pair
 * kernel and the serial scatter of Engine/mom25d/_fast.py:82-158
 * (group_vals, scatter), same loop and summation order, OpenMP over the
 * source triangles of a group (the reference's prange).  Complex numbers
 * are interleaved (re, im) doubles. */
#include <math.h>
#include <stdint.h>

typedef struct { double re, im; } cz;

static inline cz cadd(cz a, cz b) { cz r = {a.re + b.re, a.im + b.im}; return r; }
static inline cz cmul(cz a, cz b) { cz r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cz cscale(double s, cz a) { cz r = {s * a.re, s * a.im}; return r; }

static const double FOUR_PI = 12.566370614359172954;

/* scipy find_interval: x[i] <= r < x[i+1], clipped to the end intervals */
static inline int interval(const double* x, int n, double r) {
  if (r < x[1]) return 0;
  if (r >= x[n - 2]) return n - 2;
  int lo = 1, hi = n - 2;
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (r < x[mid]) hi = mid; else lo = mid;
  }
  return lo;
}

/* c: (4, n-1, 2) complex, C order */
static inline void spline2(const double* x, const cz* c, int n, double r, cz* a, cz* v) {
  int i = interval(x, n, r);
  double dx = r - x[i];
  int64_t m = n - 1;
  cz A = c[(3 * m + i) * 2 + 0], V = c[(3 * m + i) * 2 + 1];
  double z = 1.0;
  for (int k = 1; k < 4; ++k) {
    z *= dx;
    A = cadd(A, cscale(z, c[((3 - k) * m + i) * 2 + 0]));
    V = cadd(V, cscale(z, c[((3 - k) * m + i) * 2 + 1]));
  }
  *a = A;
  *v = V;
}

/* vals (npb, 3, nS, 3) complex, IVout (npb, nS) complex; sel (ns) local columns;
 * pts (ns, L, 2), w (ns, L); Vs (ns, 3, 2); cAs, cVs (ns, 3);
 * S0 (ns, M) , S1 (ns, M, 2) complex (static != 0); r (M, 2); wo (npb, K);
 * Vt (npb, 3, 2); cAt, cVt (npb, 3); spline x (n), c. */
void oracle_group_vals(int npb, int K, int nS, int ns, int L, double* vals, double* IVout, const int64_t* sel,
                       const double* pts, const double* w, const double* Vs, const double* cAs, const double* cVs,
                       int is_static, const double* S0, const double* S1, const double* r, const double* wo,
                       const double* Vt, const double* cAt, const double* cVt, const double* x, const double* c, int n,
                       double kA_re, double kA_im, double kV_re, double kV_im, double dz, double jwmu_re,
                       double jwmu_im, double ijwe_re, double ijwe_im) {
  const cz kA = {kA_re, kA_im}, kV = {kV_re, kV_im};
  const cz jwmu = {jwmu_re, jwmu_im}, ijwe = {ijwe_re, ijwe_im};
  const cz* C = (const cz*)c;
  const cz* s0 = (const cz*)S0;
  const cz* s1 = (const cz*)S1;
  cz* V = (cz*)vals;
  cz* IVo = (cz*)IVout;
  const int M = npb * K;
#pragma omp parallel for schedule(static)
  for (int s = 0; s < ns; ++s) {
    const int64_t q = sel[s];
    for (int p = 0; p < npb; ++p) {
      cz IA[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) IA[i][j].re = IA[i][j].im = 0.0;
      cz IV = {0.0, 0.0};
      for (int k = 0; k < K; ++k) {
        const int m = p * K + k;
        const double rx = r[2 * m], ry = r[2 * m + 1];
        cz k0a = {0, 0}, k1x = {0, 0}, k1y = {0, 0}, k0v = {0, 0};
        if (is_static) {
          k0a = cmul(kA, s0[(int64_t)s * M + m]);
          k1x = cmul(kA, s1[((int64_t)s * M + m) * 2]);
          k1y = cmul(kA, s1[((int64_t)s * M + m) * 2 + 1]);
          k0v = cmul(kV, s0[(int64_t)s * M + m]);
        }
        for (int l = 0; l < L; ++l) {
          const double px = pts[((int64_t)s * L + l) * 2], py = pts[((int64_t)s * L + l) * 2 + 1];
          const double ddx = rx - px, ddy = ry - py;
          const double R = sqrt(ddx * ddx + ddy * ddy);
          cz sa, sv;
          spline2(x, C, n, R, &sa, &sv);
          if (!is_static) {
            double inv = dz == 0.0 ? 1.0 / (FOUR_PI * R) : 1.0 / (FOUR_PI * sqrt(R * R + dz * dz));
            sa = cadd(sa, cscale(inv, kA));
            sv = cadd(sv, cscale(inv, kV));
          }
          const double wl = w[(int64_t)s * L + l];
          const cz saw = cscale(wl, sa);
          k0a = cadd(k0a, saw);
          k1x = cadd(k1x, cscale(px, saw));
          k1y = cadd(k1y, cscale(py, saw));
          k0v = cadd(k0v, cscale(wl, sv));
        }
        const double wk = wo[(int64_t)p * K + k];
        IV = cadd(IV, cscale(wk, k0v));
        for (int i = 0; i < 3; ++i) {
          const double ax = wk * (rx - Vt[(p * 3 + i) * 2]);
          const double ay = wk * (ry - Vt[(p * 3 + i) * 2 + 1]);
          const cz t1 = cadd(cscale(ax, k1x), cscale(ay, k1y));
          for (int j = 0; j < 3; ++j) {
            const double bb = ax * Vs[((int64_t)s * 3 + j) * 2] + ay * Vs[((int64_t)s * 3 + j) * 2 + 1];
            IA[i][j] = cadd(IA[i][j], cadd(t1, cscale(-bb, k0a)));
          }
        }
      }
      IVo[(int64_t)p * nS + q] = IV;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          cz a = cmul(cscale(cAt[p * 3 + i] * cAs[(int64_t)s * 3 + j], jwmu), IA[i][j]);
          cz b = cmul(cscale(cVt[p * 3 + i] * cVs[(int64_t)s * 3 + j], ijwe), IV);
          V[(((int64_t)p * 3 + i) * nS + q) * 3 + j] = cadd(a, b);
        }
    }
  }
}

/* Z[rows[p, i], cols[s, j]] += vals[p, i, s, j] (index < 0 skipped), serial. */
void oracle_scatter(int npb, int nS, double* Z, int64_t ldz, const double* vals, const int64_t* rows,
                    const int64_t* cols) {
  cz* Zc = (cz*)Z;
  const cz* V = (const cz*)vals;
  for (int p = 0; p < npb; ++p)
    for (int i = 0; i < 3; ++i) {
      int64_t a = rows[p * 3 + i];
      if (a < 0) continue;
      for (int s = 0; s < nS; ++s)
        for (int j = 0; j < 3; ++j) {
          int64_t b = cols[(int64_t)s * 3 + j];
          if (b >= 0) Zc[a * ldz + b] = cadd(Zc[a * ldz + b], V[(((int64_t)p * 3 + i) * nS + s) * 3 + j]);
        }
    }
}
