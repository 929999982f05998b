// datain/ reader and connectivity packing (host side); see datain.h.
#include "datain.h"

#include <fstream>

namespace dglm {

namespace {

// Size of a file in bytes, or -1 when it cannot be opened.
long long file_size(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.is_open()) return -1;
  return (long long)f.tellg();
}

template <typename T>
int read_exact(const std::string &path, std::vector<T> &v, size_t n, std::string &err) {
  const long long bytes = file_size(path);
  if (bytes < 0) {
    err = "cannot open " + path;
    return HOST_ERR_IO;
  }
  if ((size_t)bytes != n * sizeof(T)) {
    err = path + ": expected " + std::to_string(n * sizeof(T)) + " bytes, got " + std::to_string(bytes);
    return HOST_ERR_IO;
  }
  std::ifstream f(path, std::ios::binary);
  v.resize(n);
  f.read((char *)v.data(), (std::streamsize)(n * sizeof(T)));
  if ((size_t)f.gcount() != n * sizeof(T)) {
    err = path + ": short read";
    return HOST_ERR_IO;
  }
  return HOST_OK;
}

// A floating-point file written with the launcher's dtype: 'f8' (REAL =
// double) or 'f4' (REAL = float, dglm_3d_launcher.py:138-141,154).  The width
// follows from the size (n values of 8 or 4 bytes) and must equal *width once
// that is set; FP32 values are widened exactly.
int read_real(const std::string &path, std::vector<double> &v, size_t n, int *width, std::string &err) {
  const long long bytes = file_size(path);
  if (bytes < 0) {
    err = "cannot open " + path;
    return HOST_ERR_IO;
  }
  const int w = (size_t)bytes == n * 8 ? 8 : ((size_t)bytes == n * 4 ? 4 : 0);
  if (w == 0) {
    err = path + ": expected " + std::to_string(n * 8) + " (f8) or " + std::to_string(n * 4) + " (f4) bytes, got " +
          std::to_string(bytes);
    return HOST_ERR_IO;
  }
  if (*width && w != *width) {
    err = path + ": an f" + std::to_string(w) + " file among f" + std::to_string(*width) + " files";
    return HOST_ERR_IO;
  }
  *width = w;
  if (w == 8) return read_exact(path, v, n, err);
  std::vector<float> s;
  if (int rc = read_exact(path, s, n, err)) return rc;
  v.assign(s.begin(), s.end());
  return HOST_OK;
}

}  // namespace

int read_datain(const std::string &dir_in, int Np, int Nfp, Datain &d, std::string &err) {
  std::string dir(dir_in);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  std::vector<int> dim;
  int rc;
  if ((rc = read_exact(dir + "dim.bin", dim, 1, err))) return rc;
  // (the partitioned layout [K, rank, comm_size, buf_size] is not a
  // single-partition work directory: its size fails the check above)
  d.K = dim[0];
  if (d.K <= 0) {
    err = dir + "dim.bin: non-positive element count " + std::to_string(d.K);
    return HOST_ERR_IO;
  }
  if ((long long)d.K * Np >= (1LL << 31)) {
    err = dir + "dim.bin: K*Np does not fit 32-bit node indices";
    return HOST_ERR_IO;
  }
  const size_t K = (size_t)d.K, NFF = (size_t)4 * Nfp;
  int w = 0;
  if ((rc = read_exact(dir + "EtoB.bin", d.EtoB, K * 4, err)) ||
      (rc = read_exact(dir + "vmapP.bin", d.vmapP, K * NFF, err)) ||
      (rc = read_exact(dir + "vmapF.bin", d.vmapF, K * 4, err)) ||
      (rc = read_exact(dir + "Fmask.bin", d.Fmask, NFF, err)))
    return rc;
  struct {
    const char *name;
    std::vector<double> *v;
    size_t n;
  } reals[] = {{"WDr", &d.WDr, (size_t)Np * Np}, {"WDs", &d.WDs, (size_t)Np * Np}, {"WDt", &d.WDt, (size_t)Np * Np},
               {"LIFT", &d.LIFT, (size_t)Np * NFF}, {"Fscale", &d.Fscale, K * 4}, {"nx", &d.nx, K * 4},
               {"ny", &d.ny, K * 4}, {"nz", &d.nz, K * 4}, {"px", &d.px, K * Np}, {"py", &d.py, K * Np},
               {"pz", &d.pz, K * Np}, {"rx", &d.metric[0], K}, {"ry", &d.metric[1], K}, {"rz", &d.metric[2], K},
               {"sx", &d.metric[3], K}, {"sy", &d.metric[4], K}, {"sz", &d.metric[5], K}, {"tx", &d.metric[6], K},
               {"ty", &d.metric[7], K}, {"tz", &d.metric[8], K}};
  for (auto &r : reals)
    if ((rc = read_real(dir + r.name + ".bin", *r.v, r.n, &w, err))) return rc;
  d.real_bytes = w;
  return HOST_OK;
}

int pack_connectivity(int K, int Np, int Nfp, const int *vmapP, const int *vmapF, const int *EtoB, const int *fmx,
                      std::vector<int> &conn, std::vector<int> &dmap, std::vector<int> &dface, std::string &err) {
  conn.assign((size_t)K * 4, 0);
  dmap.assign((size_t)K * 4, -1);
  dface.clear();
  for (int e = 0; e < K; e++)
    for (int f = 0; f < 4; f++) {
      const int o = e * 4 + f;
      const int vf = vmapF[o];
      if (vf < 0 || vf >= 4 * K) {
        err = "vmapF of face " + std::to_string(o) + " out of range";
        return HOST_ERR_ARG;
      }
      const int nb = vf / 4, nf = vf % 4;
      int found = -1;
      for (int cde = 0; cde < 4 && found < 0; cde++) {
        bool ok = true;
        for (int j = 0; j < Nfp && ok; j++)
          ok = (long long)vmapP[(size_t)o * Nfp + j] == (long long)nb * Np + fmx[(nf * 4 + cde) * Nfp + j];
        if (ok) found = cde;
      }
      if (found < 0) {
        err = "vmapP of face " + std::to_string(o) + " is not a rotation of its neighbour face";
        return HOST_ERR_UNSUPPORTED;
      }
      if (EtoB && EtoB[o] == 1) {
        if (found != 3 || nb != e || nf != f) {
          err = "Dirichlet face " + std::to_string(o) + " (EtoB == 1) is not a boundary face (vmapP must be vmapM)";
          return HOST_ERR_ARG;
        }
        dmap[o] = (int)dface.size();
        dface.push_back(o);
      }
      conn[o] = (nb << 4) | (nf << 2) | found;
    }
  return HOST_OK;
}

}  // namespace dglm
