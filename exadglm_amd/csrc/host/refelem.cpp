// Reference-tetrahedron tables for the nodal DG operators.
//
// Computes, in extended precision (x87 long double, 64-bit mantissa) and
// rounds once to double:
//   warp-and-blend nodes (xyz on the equilateral tet, rst on the reference tet),
//   the orthonormal Vandermonde V, mass matrix M = V^-T V^-1,
//   Dr/Ds/Dt = V_{r,s,t} V^-1, the weak operators WD = M^-1 D^T M,
//   the face masks Fmask / FmaskP, the face mass matrix Emat and LIFT = V V^T Emat,
//   and the order-N -> order-(N+interpN) interpolation matrix plus the
//   high-order mass matrix used by the L2 error.
//
// Follows the algorithm of Hesthaven & Warburton, "Nodal Discontinuous
// Galerkin Methods" (2008), in the form the reference uses:
//   dim3/src/python/jacobi.py:16-114      (Gauss/GLL nodes, orthonormal Jacobi)
//   dim3/src/python/tetra.py:27-273       (warp & blend nodes, xyz->rst->abc)
//   dim3/src/python/operator_3d_ref.py:21-262 (V, M, Dr, WDr, interpM)
//   dim3/src/python/operator_3d_face.py:22-257 (Fmask, FmaskP, Emat, LIFT)
// The reference evaluates these in mpmath at 21 digits and stores doubles;
// 64-bit-mantissa arithmetic reproduces those doubles to within an ulp.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "refelem.h"

namespace dglm {

typedef long double R;
typedef std::vector<R> Vec;

namespace {

const R PI_L = 3.141592653589793238462643383279502884L;

// Dense row-major matrix in extended precision.
struct Mat {
  int n = 0, m = 0;
  Vec a;
  Mat() {}
  Mat(int n_, int m_) : n(n_), m(m_), a((size_t)n_ * m_, 0.0L) {}
  R &operator()(int i, int j) { return a[(size_t)i * m + j]; }
  R operator()(int i, int j) const { return a[(size_t)i * m + j]; }
};

Mat matmul(const Mat &A, const Mat &B) {
  Mat C(A.n, B.m);
  for (int i = 0; i < A.n; i++)
    for (int k = 0; k < A.m; k++) {
      R aik = A(i, k);
      if (aik == 0) continue;
      for (int j = 0; j < B.m; j++) C(i, j) += aik * B(k, j);
    }
  return C;
}

Mat transpose(const Mat &A) {
  Mat T(A.m, A.n);
  for (int i = 0; i < A.n; i++)
    for (int j = 0; j < A.m; j++) T(j, i) = A(i, j);
  return T;
}

// Gauss-Jordan inverse with partial pivoting.
Mat inverse(const Mat &A) {
  int n = A.n;
  Mat W(n, 2 * n);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) W(i, j) = A(i, j);
    W(i, n + i) = 1;
  }
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabsl(W(r, c)) > fabsl(W(p, c))) p = r;
    if (W(p, c) == 0) throw std::runtime_error("singular matrix in operator generation");
    if (p != c)
      for (int j = 0; j < 2 * n; j++) std::swap(W(p, j), W(c, j));
    R inv = 1 / W(c, c);
    for (int j = 0; j < 2 * n; j++) W(c, j) *= inv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      R f = W(r, c);
      if (f == 0) continue;
      for (int j = 0; j < 2 * n; j++) W(r, j) -= f * W(c, j);
    }
  }
  Mat I(n, n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) I(i, j) = W(i, n + j);
  return I;
}

R ipow(R x, int p) {
  R r = 1;
  for (int i = 0; i < p; i++) r *= x;
  return r;
}

// Orthonormal Jacobi polynomial P_n^{(a,b)}(x) by the three-term recurrence
// (jacobi.py:69-101).
R jacobi_p(int n, R a, R b, R x) {
  R gamma0 = powl(2.0L, a + b + 1) / (a + b + 1) * tgammal(a + 1) * tgammal(b + 1) / tgammal(a + b + 1);
  R p0 = 1 / sqrtl(gamma0);
  if (n == 0) return p0;
  R gamma1 = (a + 1) * (b + 1) / (a + b + 3) * gamma0;
  R p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / sqrtl(gamma1);
  if (n == 1) return p1;
  R aold = 2 / (a + b + 2) * sqrtl((a + 1) * (b + 1) / (a + b + 3));
  R pm1 = p0, pc = p1;
  for (int i = 1; i < n; i++) {
    R h1 = 2 * i + a + b;
    R anew = 2 / (h1 + 2) * sqrtl((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3));
    R bnew = -(a * a - b * b) / h1 / (h1 + 2);
    R pn = 1 / anew * (-aold * pm1 + (x - bnew) * pc);
    pm1 = pc;
    pc = pn;
    aold = anew;
  }
  return pc;
}

// d/dx P_n^{(a,b)} = sqrt(n(n+a+b+1)) P_{n-1}^{(a+1,b+1)} (jacobi.py:104-114).
R grad_jacobi_p(int n, R a, R b, R x) {
  if (n == 0) return 0;
  return sqrtl((R)n * (n + a + b + 1)) * jacobi_p(n - 1, a + 1, b + 1, x);
}

// Zeros of P_n^{(a,b)}, ascending, by Newton iteration with deflation from
// Chebyshev-Gauss initial guesses.  For odd n the middle zero of the
// symmetric (a == b) family is set to exactly 0 (jacobi.py:49-50).
Vec jacobi_zeros(int n, R a, R b) {
  Vec x(n);
  for (int k = 0; k < n; k++) {
    R r = -cosl((2 * k + 1) * PI_L / (2 * n));
    if (k > 0) r = (r + x[k - 1]) / 2;
    for (int it = 0; it < 200; it++) {
      R s = 0;
      for (int i = 0; i < k; i++) s += 1 / (r - x[i]);
      R p = jacobi_p(n, a, b, r);
      R dp = grad_jacobi_p(n, a, b, r);
      R delta = -p / (dp - s * p);
      r += delta;
      if (fabsl(delta) < 1e-30L) break;
    }
    x[k] = r;
  }
  std::sort(x.begin(), x.end());
  if (n % 2 == 1 && a == b) x[n / 2] = 0;
  return x;
}

}  // namespace

// Gauss-Lobatto-Legendre nodes on [-1,1], ascending (jacobi.py:53-66).
std::vector<long double> gll_nodes(int N) {
  Vec x(N + 1);
  x[0] = -1;
  x[N] = 1;
  if (N > 1) {
    Vec in = jacobi_zeros(N - 1, 1, 1);
    for (int i = 0; i < N - 1; i++) x[i + 1] = in[i];
  }
  return x;
}

long double jacobi_value(int n, long double a, long double b, long double x) { return jacobi_p(n, a, b, x); }
long double jacobi_grad(int n, long double a, long double b, long double x) { return grad_jacobi_p(n, a, b, x); }

namespace {

// Warp function along an edge: interpolant of (GLL - equidistant) node
// displacements, divided by the (1-r^2) blend (tetra.py:27-49).
Vec eval_warp(int N, const Vec &xnodes, const Vec &xout) {
  int np = (int)xout.size();
  Vec warp(np, 0), d(np);
  Vec xeq(N + 1);
  // The equidistant nodes are formed in double precision as the reference
  // does (a Python float expression), then promoted.
  for (int i = 1; i <= N + 1; i++) xeq[i - 1] = (R)(-1.0 + (double)(2 * (N + 1 - i)) / N);
  for (int i = 1; i <= N + 1; i++) {
    for (int k = 0; k < np; k++) d[k] = xnodes[i - 1] - xeq[i - 1];
    for (int j = 2; j <= N; j++)
      if (i != j)
        for (int k = 0; k < np; k++) d[k] = d[k] * (xout[k] - xeq[j - 1]) / (xeq[i - 1] - xeq[j - 1]);
    if (i != 1)
      for (int k = 0; k < np; k++) d[k] = -d[k] / (xeq[i - 1] - xeq[0]);
    if (i != N + 1)
      for (int k = 0; k < np; k++) d[k] = d[k] / (xeq[i - 1] - xeq[N]);
    for (int k = 0; k < np; k++) warp[k] += d[k];
  }
  return warp;
}

// Two-dimensional warp & blend displacement on one face (tetra.py:52-100).
void warp_shift(int N, R pval, const Vec &L1, const Vec &L2, const Vec &L3, Vec &dx, Vec &dy) {
  int np = (int)L1.size();
  Vec g = gll_nodes(N);
  Vec gaussX(N + 1);
  for (int i = 0; i <= N; i++) gaussX[i] = -g[i];
  Vec a32(np), a13(np), a21(np);
  for (int k = 0; k < np; k++) {
    a32[k] = L3[k] - L2[k];
    a13[k] = L1[k] - L3[k];
    a21[k] = L2[k] - L1[k];
  }
  Vec wf1 = eval_warp(N, gaussX, a32), wf2 = eval_warp(N, gaussX, a13), wf3 = eval_warp(N, gaussX, a21);
  dx.assign(np, 0);
  dy.assign(np, 0);
  const R c2 = cosl(2 * PI_L / 3), c4 = cosl(4 * PI_L / 3), s2 = sinl(2 * PI_L / 3), s4 = sinl(4 * PI_L / 3);
  for (int k = 0; k < np; k++) {
    R w1 = L2[k] * L3[k] * (4 * wf1[k]) * (1 + (pval * L1[k]) * (pval * L1[k]));
    R w2 = L1[k] * L3[k] * (4 * wf2[k]) * (1 + (pval * L2[k]) * (pval * L2[k]));
    R w3 = L1[k] * L2[k] * (4 * wf3[k]) * (1 + (pval * L3[k]) * (pval * L3[k]));
    dx[k] = 1 * w1 + c2 * w2 + c4 * w3;
    dy[k] = 0 * w1 + s2 * w2 + s4 * w3;
  }
}

struct Verts {
  R v[4][3];
  Verts() {
    R s3 = sqrtl(3.0L), s6 = sqrtl(6.0L);
    R t[4][3] = {{-1, -1 / s3, -1 / s6}, {1, -1 / s3, -1 / s6}, {0, 2 / s3, -1 / s6}, {0, 0, 3 / s6}};
    memcpy(v, t, sizeof(t));
  }
};

}  // namespace

int tet_np(int N) { return (N + 1) * (N + 2) * (N + 3) / 6; }
int tri_np(int N) { return (N + 1) * (N + 2) / 2; }

// Warp & blend nodes on the equilateral tetrahedron (tetra.py:103-225).
void tet_nodes_xyz(int N, std::vector<long double> &xyz) {
  static const double alpha_store[15] = {0, 0, 0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                                         1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655};
  const R TOL = 1e-15L;
  R alpha = N <= 15 ? (R)alpha_store[N - 1] : 1.0L;
  int np = tet_np(N);
  Vec r(np), s(np), t(np);
  int sk = 0;
  for (int n = 1; n <= N + 1; n++)
    for (int m = 1; m <= N + 2 - n; m++)
      for (int q = 1; q <= N + 3 - n - m; q++) {
        // Equidistant nodes formed in double precision as in the reference.
        r[sk] = (R)(-1.0 + (double)((q - 1) * 2) / N);
        s[sk] = (R)(-1.0 + (double)((m - 1) * 2) / N);
        t[sk] = (R)(-1.0 + (double)((n - 1) * 2) / N);
        sk++;
      }
  Vec L1(np), L2(np), L3(np), L4(np);
  for (int k = 0; k < np; k++) {
    L1[k] = (1 + t[k]) / 2;
    L2[k] = (1 + s[k]) / 2;
    L3[k] = -(1 + r[k] + s[k] + t[k]) / 2;
    L4[k] = (1 + r[k]) / 2;
  }
  Verts V;
  R t1[4][3], t2[4][3];
  for (int k = 0; k < 3; k++) {
    t1[0][k] = V.v[1][k] - V.v[0][k];
    t1[1][k] = V.v[1][k] - V.v[0][k];
    t1[2][k] = V.v[2][k] - V.v[1][k];
    t1[3][k] = V.v[2][k] - V.v[0][k];
    t2[0][k] = V.v[2][k] - 0.5L * (V.v[0][k] + V.v[1][k]);
    t2[1][k] = V.v[3][k] - 0.5L * (V.v[0][k] + V.v[1][k]);
    t2[2][k] = V.v[3][k] - 0.5L * (V.v[1][k] + V.v[2][k]);
    t2[3][k] = V.v[3][k] - 0.5L * (V.v[0][k] + V.v[2][k]);
  }
  for (int n = 0; n < 4; n++) {
    R n1 = sqrtl(t1[n][0] * t1[n][0] + t1[n][1] * t1[n][1] + t1[n][2] * t1[n][2]);
    R n2 = sqrtl(t2[n][0] * t2[n][0] + t2[n][1] * t2[n][1] + t2[n][2] * t2[n][2]);
    for (int k = 0; k < 3; k++) {
      t1[n][k] /= n1;
      t2[n][k] /= n2;
    }
  }
  xyz.assign((size_t)np * 3, 0);
  Vec shift((size_t)np * 3, 0);
  for (int k = 0; k < np; k++)
    for (int j = 0; j < 3; j++)
      xyz[k * 3 + j] = L3[k] * V.v[0][j] + L4[k] * V.v[1][j] + L2[k] * V.v[2][j] + L1[k] * V.v[3][j];
  for (int face = 0; face < 4; face++) {
    const Vec *La, *Lb, *Lc, *Ld;
    if (face == 0) { La = &L1; Lb = &L2; Lc = &L3; Ld = &L4; }
    else if (face == 1) { La = &L2; Lb = &L1; Lc = &L3; Ld = &L4; }
    else if (face == 2) { La = &L3; Lb = &L1; Lc = &L4; Ld = &L2; }
    else { La = &L4; Lb = &L1; Lc = &L3; Ld = &L2; }
    Vec w1, w2;
    warp_shift(N, alpha, *Lb, *Lc, *Ld, w1, w2);
    for (int k = 0; k < np; k++) {
      R la = (*La)[k], lb = (*Lb)[k], lc = (*Lc)[k], ld = (*Ld)[k];
      R blend = lb * lc * ld;
      R denom = (lb + 0.5L * la) * (lc + 0.5L * la) * (ld + 0.5L * la);
      if (denom > TOL) blend = (1 + (alpha * la) * (alpha * la)) * blend / denom;
      bool on_face = la < TOL && ((lb > TOL) + (lc > TOL) + (ld > TOL)) < 3;
      for (int j = 0; j < 3; j++) {
        if (on_face)
          shift[k * 3 + j] = w1[k] * t1[face][j] + w2[k] * t2[face][j];
        else
          shift[k * 3 + j] += blend * w1[k] * t1[face][j] + blend * w2[k] * t2[face][j];
      }
    }
  }
  for (size_t i = 0; i < xyz.size(); i++) xyz[i] += shift[i];
}

// Equilateral (x,y,z) -> reference (r,s,t) by the affine vertex map (tetra.py:228-246).
void xyz_to_rst(const std::vector<long double> &xyz, std::vector<long double> &rst) {
  Verts V;
  int np = (int)xyz.size() / 3;
  Mat A(3, 3);
  for (int i = 0; i < 3; i++) {
    A(i, 0) = 0.5L * (V.v[1][i] - V.v[0][i]);
    A(i, 1) = 0.5L * (V.v[2][i] - V.v[0][i]);
    A(i, 2) = 0.5L * (V.v[3][i] - V.v[0][i]);
  }
  Mat Ai = inverse(A);
  rst.assign((size_t)np * 3, 0);
  for (int k = 0; k < np; k++) {
    R rhs[3];
    for (int j = 0; j < 3; j++) rhs[j] = xyz[k * 3 + j] - 0.5L * (V.v[1][j] + V.v[2][j] + V.v[3][j] - V.v[0][j]);
    for (int i = 0; i < 3; i++) rst[k * 3 + i] = Ai(i, 0) * rhs[0] + Ai(i, 1) * rhs[1] + Ai(i, 2) * rhs[2];
  }
}

// Collapsed coordinates (r,s,t) -> (a,b,c) (tetra.py:249-273).
void rst_to_abc(const std::vector<long double> &rst, std::vector<long double> &abc) {
  const R TOL = 1e-15L;
  int np = (int)rst.size() / 3;
  abc.assign((size_t)np * 3, 0);
  for (int k = 0; k < np; k++) {
    R r = rst[k * 3], s = rst[k * 3 + 1], t = rst[k * 3 + 2];
    abc[k * 3 + 0] = fabsl(s + t) > TOL ? -2 * (1 + r) / (s + t) - 1 : -1;
    abc[k * 3 + 1] = fabsl(t - 1) > TOL ? 2 * (1 + s) / (1 - t) - 1 : -1;
    abc[k * 3 + 2] = t;
  }
}

namespace {

// Orthonormal modal basis on the tetrahedron (operator_3d_ref.py:21-42).
R simplex3(int i, int j, int k, R a, R b, R c) {
  R h1 = jacobi_p(i, 0, 0, a);
  R h2 = jacobi_p(j, 2 * i + 1, 0, b);
  R h3 = jacobi_p(k, 2 * (i + j) + 2, 0, c);
  return 2 * sqrtl(2.0L) * h1 * h2 * ipow(1 - b, i) * h3 * ipow(1 - c, i + j);
}

// Gradient of the modal basis with respect to (r,s,t) (operator_3d_ref.py:45-112).
void grad_simplex3(int i, int j, int k, R a, R b, R c, R &dr, R &ds, R &dt) {
  R fa = jacobi_p(i, 0, 0, a), dfa = grad_jacobi_p(i, 0, 0, a);
  R gb = jacobi_p(j, 2 * i + 1, 0, b), dgb = grad_jacobi_p(j, 2 * i + 1, 0, b);
  R hc = jacobi_p(k, 2 * (i + j) + 2, 0, c), dhc = grad_jacobi_p(k, 2 * (i + j) + 2, 0, c);
  R hb = 0.5L * (1 - b), hcc = 0.5L * (1 - c);
  R vr = dfa * gb * hc;
  if (i > 0) vr *= ipow(hb, i - 1);
  if (i + j > 0) vr *= ipow(hcc, i + j - 1);
  R vs = 0.5L * (1 + a) * vr;
  R tmp = dgb * ipow(hb, i);
  if (i > 0) tmp = tmp + (-0.5L * i) * (gb * ipow(hb, i - 1));
  if (i + j > 0) tmp = tmp * ipow(hcc, i + j - 1);
  tmp = fa * tmp * hc;
  vs += tmp;
  R vt = 0.5L * (1 + a) * vr + 0.5L * (1 + b) * tmp;
  tmp = dhc * ipow(hcc, i + j);
  if (i + j > 0) tmp -= 0.5L * (i + j) * hc * ipow(hcc, i + j - 1);
  tmp = fa * gb * tmp * ipow(hb, i);
  vt += tmp;
  R scale = powl(2.0L, 2 * i + j + 1.5L);
  dr = scale * vr;
  ds = scale * vs;
  dt = scale * vt;
}

Mat vandermonde3(int N, const Vec &abc) {
  int np = (int)abc.size() / 3, nm = tet_np(N);
  Mat V(np, nm);
  int sk = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N - i; j++)
      for (int k = 0; k <= N - i - j; k++) {
        for (int q = 0; q < np; q++) V(q, sk) = simplex3(i, j, k, abc[q * 3], abc[q * 3 + 1], abc[q * 3 + 2]);
        sk++;
      }
  return V;
}

// Orthonormal modal basis on the triangle (operator_3d_face.py:126-142).
R simplex2(int i, int j, R a, R b) {
  return sqrtl(2.0L) * jacobi_p(i, 0, 0, a) * jacobi_p(j, 2 * i + 1, 0, b) * ipow(1 - b, i);
}

void to_double(const Mat &A, double *out) {
  if (!out) return;
  for (size_t i = 0; i < A.a.size(); i++) out[i] = (double)A.a[i];
}

// Nested-list face-node manipulations of operator_3d_face.py:28-71: a face's
// node list is split into rows of N+1, N, ..., 1 nodes; "rotate" reads the
// rows bottom-up column by column, "invert" reverses each row.
typedef std::vector<std::vector<int>> Nested;

Nested to_nested(int N, const std::vector<int> &v) {
  Nested out;
  size_t k = 0;
  for (int n = N + 1; n > 0; n--) {
    std::vector<int> row;
    for (int q = 0; q < n; q++) row.push_back(v[k++]);
    out.push_back(row);
  }
  return out;
}

std::vector<int> flatten(const Nested &nl) {
  std::vector<int> out;
  for (auto &r : nl) out.insert(out.end(), r.begin(), r.end());
  return out;
}

Nested rotate_rows(Nested nl) {
  Nested out;
  while (!nl[0].empty()) {
    std::vector<int> row;
    for (int q = (int)nl.size() - 1; q >= 0; q--) {
      if (nl[q].empty()) continue;
      row.push_back(nl[q].front());
      nl[q].erase(nl[q].begin());
    }
    out.push_back(row);
  }
  return out;
}

Nested invert_rows(const Nested &nl) {
  Nested out;
  for (auto &r : nl) out.push_back(std::vector<int>(r.rbegin(), r.rend()));
  return out;
}

}  // namespace

void face_masks(int N, const std::vector<long double> &rst, int *Fmask, int *FmaskP) {
  const R NODETOL = 1e-12L;
  int np = tet_np(N), nfp = tri_np(N);
  for (int fi = 0; fi < 4; fi++) {
    std::vector<int> face;
    for (int k = 0; k < np; k++) {
      R r = rst[k * 3], s = rst[k * 3 + 1], t = rst[k * 3 + 2];
      R g = fi == 0 ? t + 1 : fi == 1 ? s + 1 : fi == 2 ? r + s + t + 1 : r + 1;
      if (fabsl(g) < NODETOL) face.push_back(k);
    }
    if ((int)face.size() != nfp) throw std::runtime_error("face node classification failed");
    Nested f = to_nested(N, face);
    if (fi == 0 || fi == 3) f = invert_rows(f);
    Nested r1 = rotate_rows(f), r2 = rotate_rows(r1);
    std::vector<int> r0 = flatten(f), i0 = flatten(invert_rows(f)), i1 = flatten(invert_rows(r1)),
                     i2 = flatten(invert_rows(r2));
    for (int j = 0; j < nfp; j++) {
      Fmask[fi * nfp + j] = r0[j];
      FmaskP[(fi * 3 + 0) * nfp + j] = i0[j];
      FmaskP[(fi * 3 + 1) * nfp + j] = i1[j];
      FmaskP[(fi * 3 + 2) * nfp + j] = i2[j];
    }
  }
}

void reference_operators(int N, RefOps &out) {
  if (N < 1 || N > 10) throw std::runtime_error("polynomial order must be in [1, 10]");
  int np = tet_np(N), nfp = tri_np(N), nff = 4 * nfp;
  Vec xyz, rst, abc;
  tet_nodes_xyz(N, xyz);
  xyz_to_rst(xyz, rst);
  rst_to_abc(rst, abc);

  Mat V = vandermonde3(N, abc);
  Mat Vi = inverse(V);
  Mat M = matmul(transpose(Vi), Vi);
  Mat Mi = matmul(V, transpose(V));
  Mat Vr(np, np), Vs(np, np), Vt(np, np);
  int sk = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N - i; j++)
      for (int k = 0; k <= N - i - j; k++) {
        for (int q = 0; q < np; q++)
          grad_simplex3(i, j, k, abc[q * 3], abc[q * 3 + 1], abc[q * 3 + 2], Vr(q, sk), Vs(q, sk), Vt(q, sk));
        sk++;
      }
  Mat Dr = matmul(Vr, Vi), Ds = matmul(Vs, Vi), Dt = matmul(Vt, Vi);
  Mat WDr = matmul(matmul(Mi, transpose(Dr)), M);
  Mat WDs = matmul(matmul(Mi, transpose(Ds)), M);
  Mat WDt = matmul(matmul(Mi, transpose(Dt)), M);

  out.N = N;
  out.Np = np;
  out.Nfp = nfp;
  out.Fmask.assign(4 * nfp, 0);
  out.FmaskP.assign(12 * nfp, 0);
  face_masks(N, rst, out.Fmask.data(), out.FmaskP.data());

  // Face mass matrices in Fmask order; each face uses the pair of collapsed
  // coordinates that parametrise it (operator_3d_face.py:168-252).
  static const int coord_pair[4][2] = {{0, 1}, {0, 2}, {1, 2}, {1, 2}};
  Mat Emat(np, nff);
  for (int f = 0; f < 4; f++) {
    Mat V2(nfp, nfp);
    int c0 = coord_pair[f][0], c1 = coord_pair[f][1];
    int q2 = 0;
    for (int i = 0; i <= N; i++)
      for (int j = 0; j <= N - i; j++) {
        for (int q = 0; q < nfp; q++) {
          int node = out.Fmask[f * nfp + q];
          V2(q, q2) = simplex2(i, j, abc[node * 3 + c0], abc[node * 3 + c1]);
        }
        q2++;
      }
    Mat V2i = inverse(V2);
    Mat E = matmul(transpose(V2i), V2i);
    for (int i = 0; i < nfp; i++)
      for (int j = 0; j < nfp; j++) Emat(out.Fmask[f * nfp + i], f * nfp + j) = E(i, j);
  }
  Mat LIFT = matmul(V, matmul(transpose(V), Emat));

  auto fill = [](const Mat &A, std::vector<double> &v) {
    v.resize(A.a.size());
    to_double(A, v.data());
  };
  std::vector<double> tmp;
  out.xyz.resize(xyz.size());
  out.rst.resize(rst.size());
  for (size_t i = 0; i < xyz.size(); i++) out.xyz[i] = (double)xyz[i];
  for (size_t i = 0; i < rst.size(); i++) out.rst[i] = (double)rst[i];
  fill(M, out.M);
  fill(Mi, out.invM);
  fill(V, out.V);
  fill(Dr, out.Dr);
  fill(Ds, out.Ds);
  fill(Dt, out.Dt);
  fill(WDr, out.WDr);
  fill(WDs, out.WDs);
  fill(WDt, out.WDt);
  fill(Emat, out.Emat);
  fill(LIFT, out.LIFT);
}

void interp_operators(int N, int high_N, std::vector<double> &high_rst, std::vector<double> &interpM,
                      std::vector<double> &high_M) {
  if (high_N < N || high_N > 14) throw std::runtime_error("interpolation order out of range");
  Vec xyz, rst, abc, hxyz, hrst, habc;
  tet_nodes_xyz(N, xyz);
  xyz_to_rst(xyz, rst);
  rst_to_abc(rst, abc);
  tet_nodes_xyz(high_N, hxyz);
  xyz_to_rst(hxyz, hrst);
  rst_to_abc(hrst, habc);
  Mat Vi = inverse(vandermonde3(N, abc));
  Mat IV = vandermonde3(N, habc);  // high_Np x Np: order-N basis at high nodes
  Mat IM = matmul(IV, Vi);
  Mat HV = vandermonde3(high_N, habc);
  Mat HVi = inverse(HV);
  Mat HM = matmul(transpose(HVi), HVi);
  high_rst.resize(hrst.size());
  for (size_t i = 0; i < hrst.size(); i++) high_rst[i] = (double)hrst[i];
  interpM.resize(IM.a.size());
  to_double(IM, interpM.data());
  high_M.resize(HM.a.size());
  to_double(HM, high_M.data());
}

}  // namespace dglm
