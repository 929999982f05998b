// Self-check of the host-side C++ of the library (operator tables, uniform
// mesh, affine geometry, vmaps, the datain/ reader and the connectivity
// packing), for a build with AddressSanitizer and UndefinedBehaviorSanitizer:
// the counterpart of the reference launcher's build(use_sanitize=True), which
// compiles its CPU solver with -fsanitize=address,undefined
// (dim3/src/python/dglm_3d_launcher.py:389-390).  The GPU kernels are not
// part of it (no sanitizer runs device code here).
//
//   selfcheck              operators for N = 1..10 and the mesh pipeline on a
//                          periodic and a tagged 4^3-vertex box, N = 1..4
//   selfcheck DIR N        read the datain directory DIR for order N and pack
//                          its connectivity
//
// Exit status 0: every check passed; 1: a check failed; 2: DIR is not a
// readable datain directory for order N (the message says why).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "datain.h"
#include "mesh.h"
#include "refelem.h"

using namespace dglm;

namespace {

int g_fail = 0;

void check(bool ok, const std::string &what) {
  if (!ok) {
    std::fprintf(stderr, "selfcheck: FAILED %s\n", what.c_str());
    g_fail = 1;
  }
}

// Face-node table of the packing: FmaskP[f][r] at code r < 3, Fmask[f] at 3.
std::vector<int> face_table(const RefOps &r) {
  std::vector<int> fmx(16 * r.Nfp);
  for (int f = 0; f < 4; f++)
    for (int c = 0; c < 4; c++)
      for (int j = 0; j < r.Nfp; j++)
        fmx[(f * 4 + c) * r.Nfp + j] = c < 3 ? r.FmaskP[(f * 3 + c) * r.Nfp + j] : r.Fmask[f * r.Nfp + j];
  return fmx;
}

void check_operators(int N) {
  RefOps r;
  reference_operators(N, r);
  const int Np = r.Np, Nfp = r.Nfp;
  check(Np == tet_np(N) && Nfp == tri_np(N), "node counts N=" + std::to_string(N));
  double msum = 0.0, esum = 0.0;
  for (double x : r.M) msum += x;
  for (double x : r.Emat) esum += x;
  check(std::fabs(msum - 4.0 / 3.0) < 1e-12, "sum M = 4/3 at N=" + std::to_string(N));
  check(std::fabs(esum - 8.0) < 1e-11, "sum Emat = 8 at N=" + std::to_string(N));
  for (int i = 0; i < Np; i++) {
    double s = 0.0;
    for (int j = 0; j < Np; j++) s += r.Dr[(size_t)i * Np + j];
    check(std::fabs(s) < 1e-9, "Dr row sum at N=" + std::to_string(N));
  }
  std::vector<double> hr, im, hm;
  interp_operators(N, N + 2, hr, im, hm);
  check(im.size() == (size_t)tet_np(N + 2) * Np, "interpolation shape at N=" + std::to_string(N));
}

void check_mesh(int N, const int tags[3]) {
  RefOps r;
  reference_operators(N, r);
  const int n = 4, K = 6 * (n - 1) * (n - 1) * (n - 1), Np = r.Np, Nfp = r.Nfp;
  std::vector<double> V((size_t)n * n * n * 3);
  std::vector<int> EtoV(K * 4), EtoE(K * 4), EtoF(K * 4), EtoR(K * 4), EtoB(K * 4);
  uniform_vertices(n, n, n, 0.0, 1.0, 0.0, 2.0, -1.0, 1.0, V.data());
  uniform_connectivity(n, n, n, tags, EtoV.data(), EtoE.data(), EtoF.data(), EtoR.data(), EtoB.data());
  std::vector<double> PX((size_t)K * Np), PY(PX.size()), PZ(PX.size()), J(K), mt(9 * (size_t)K);
  element_points(K, Np, EtoV.data(), V.data(), r.rst.data(), PX.data(), PY.data(), PZ.data());
  double *m9[9];
  for (int q = 0; q < 9; q++) m9[q] = &mt[(size_t)q * K];
  element_metrics(K, EtoV.data(), V.data(), J.data(), m9);
  double vol = 0.0;
  for (double j : J) vol += j * 4.0 / 3.0;
  check(std::fabs(vol - 4.0) < 1e-12, "sum of element volumes");
  std::vector<double> nx(K * 4), ny(K * 4), nz(K * 4), sJ(K * 4), Fs(K * 4), h(K);
  face_normals(K, J.data(), m9, nx.data(), ny.data(), nz.data(), sJ.data(), Fs.data());
  for (int o = 0; o < K * 4; o++)
    check(std::fabs(nx[o] * nx[o] + ny[o] * ny[o] + nz[o] * nz[o] - 1.0) < 1e-13, "unit normals");
  std::vector<int> vM((size_t)K * 4 * Nfp), vP(vM.size()), vF(K * 4);
  face_vmaps(K, Np, Nfp, EtoE.data(), EtoF.data(), EtoR.data(), r.Fmask.data(), r.FmaskP.data(), vM.data(),
             vP.data(), vF.data());
  spatial_h(K, EtoV.data(), V.data(), h.data());
  for (double x : h) check(x > 0.0, "positive spatial h");
  // Dirichlet-tagged faces exist only where an axis is tagged 1.
  const bool dir = tags[0] == 1 || tags[1] == 1 || tags[2] == 1;
  std::vector<int> conn, dmap, dface;
  std::string err;
  const std::vector<int> fmx = face_table(r);
  const int rc = pack_connectivity(K, Np, Nfp, vP.data(), vF.data(), EtoB.data(), fmx.data(), conn, dmap, dface, err);
  check(rc == HOST_OK, "pack_connectivity of a generated mesh: " + err);
  check(dir == !dface.empty(), "Dirichlet faces found exactly on tagged axes");
  // The packed form reproduces vmapP, and neighbours are an involution.
  for (int o = 0; o < K * 4 && rc == HOST_OK; o++) {
    const int c = conn[o], nb = c >> 4, nf = (c >> 2) & 3, code = c & 3;
    for (int j = 0; j < Nfp; j++) check(vP[(size_t)o * Nfp + j] == nb * Np + fmx[(nf * 4 + code) * Nfp + j], "vmapP");
    check(vF[nb * 4 + nf] == o, "neighbour involution");
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 1) {
    for (int N = 1; N <= 10; N++) check_operators(N);
    const int periodic[3] = {-1, -1, -1}, tagged[3] = {1, -1, 2};
    for (int N = 1; N <= 4; N++) {
      check_mesh(N, periodic);
      check_mesh(N, tagged);
    }
  } else if (argc == 3) {
    const int N = std::atoi(argv[2]);
    if (N < 1 || N > 14) {
      std::fprintf(stderr, "selfcheck: bad order %s\n", argv[2]);
      return 2;
    }
    Datain d;
    std::string err;
    if (int rc = read_datain(argv[1], tet_np(N), tri_np(N), d, err)) {
      std::fprintf(stderr, "selfcheck: datain error %d: %s\n", rc, err.c_str());
      return 2;
    }
    RefOps r;
    reference_operators(N, r);
    for (int i = 0; i < 4 * r.Nfp; i++)
      if (d.Fmask[i] != r.Fmask[i]) {
        std::fprintf(stderr, "selfcheck: Fmask.bin does not match order %d\n", N);
        return 2;
      }
    const std::vector<int> fmx = face_table(r);
    std::vector<int> conn, dmap, dface;
    if (int rc = pack_connectivity(d.K, r.Np, r.Nfp, d.vmapP.data(), d.vmapF.data(), d.EtoB.data(), fmx.data(), conn,
                                   dmap, dface, err)) {
      std::fprintf(stderr, "selfcheck: connectivity error %d: %s\n", rc, err.c_str());
      return 2;
    }
    std::printf("selfcheck: datain K=%d f%d, %zu Dirichlet faces\n", d.K, d.real_bytes, dface.size());
  } else {
    std::fprintf(stderr, "usage: selfcheck [DATAIN_DIR N]\n");
    return 2;
  }
  if (!g_fail) std::printf("selfcheck: ok\n");
  return g_fail;
}
