// Uniform periodic tetrahedral mesh and its affine geometry / connectivity.
//
//   uniform box -> 6 tets per hex sharing the diagonal v0-v7
//                  (dim3/src/cpp/uniform_3d_single_core.cpp:8-67,
//                   dim3/src/python/uniform_3d_single.py:43-98)
//   EtoE/EtoF/EtoR/EtoB with periodic wrap or physical tags
//                  (uniform_3d_single_core.cpp:70-277)
//   node coordinates, affine metrics, normals, sJ, Fscale, vmaps, h
//                  (dim3/src/cpp/mesh_3d_core.cpp:10-145, 237-356, 423-505)
//
// The floating-point expressions for coordinates and metrics are evaluated in
// the same order as the reference so that the mesh arrays agree bit for bit
// (this file is compiled without FP contraction).
#include <cmath>
#include <stdexcept>
#include <vector>

#include "mesh.h"

namespace dglm {

namespace {

// Local face-neighbour table of the 6-tet hex split.  For tet l and face f:
// which cube the neighbour is in, its local tet, its face and the rotation.
// Faces: 0=(v1,v0,v2) 1=(v0,v1,v3) 2=(v1,v2,v3) 3=(v2,v0,v3); the six tets
// are [v0,v7,v3,v1] [v0,v7,v2,v3] [v0,v7,v6,v2] [v0,v7,v4,v6] [v0,v7,v5,v4]
// [v0,v7,v1,v5] of hex (v0..v7) = (i..i+1, j..j+1, k..k+1), k fastest.
enum Dir { SAME = 0, XM, XP, YM, YP, ZM, ZP };
struct Nbr {
  int dir, tet, face, rot;
};
const Nbr kNbr[6][4] = {
    {{SAME, 1, 1, 0}, {SAME, 5, 0, 0}, {ZP, 2, 3, 1}, {XM, 4, 2, 1}},
    {{SAME, 2, 1, 0}, {SAME, 0, 0, 0}, {YP, 5, 3, 2}, {XM, 3, 2, 2}},
    {{SAME, 3, 1, 0}, {SAME, 1, 0, 0}, {YP, 4, 3, 1}, {ZM, 0, 2, 1}},
    {{SAME, 4, 1, 0}, {SAME, 2, 0, 0}, {XP, 1, 3, 2}, {ZM, 5, 2, 2}},
    {{SAME, 5, 1, 0}, {SAME, 3, 0, 0}, {XP, 0, 3, 1}, {YM, 2, 2, 1}},
    {{SAME, 0, 1, 0}, {SAME, 4, 0, 0}, {ZP, 3, 3, 2}, {YM, 1, 2, 2}},
};

}  // namespace

void uniform_vertices(int nx, int ny, int nz, double x1, double x2, double y1, double y2, double z1, double z2,
                      double *VXYZ) {
  // np.linspace semantics: start + i*step, last point exactly the end point.
  auto lin = [](double a, double b, int n, int i) {
    if (n == 1) return a;
    if (i == n - 1) return b;
    double step = (b - a) / (n - 1);
    return (double)i * step + a;
  };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        size_t v = ((size_t)i * ny + j) * nz + k;
        VXYZ[v * 3 + 0] = lin(x1, x2, nx, i);
        VXYZ[v * 3 + 1] = lin(y1, y2, ny, j);
        VXYZ[v * 3 + 2] = lin(z1, z2, nz, k);
      }
}

void uniform_connectivity(int nx, int ny, int nz, const int phy_tags[3], int *EtoV, int *EtoE, int *EtoF, int *EtoR,
                          int *EtoB) {
  if (nx < 2 || ny < 2 || nz < 2) throw std::runtime_error("uniform mesh needs at least 2 vertices per axis");
  const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
  auto node = [&](int i, int j, int k) { return (i * ny + j) * nz + k; };
  auto cube = [&](int i, int j, int k) { return (i * cy + j) * cz + k; };
  for (int i = 0; i < cx; i++)
    for (int j = 0; j < cy; j++)
      for (int k = 0; k < cz; k++) {
        int c = cube(i, j, k);
        int v[8];
        for (int q = 0; q < 8; q++) v[q] = node(i + ((q >> 2) & 1), j + ((q >> 1) & 1), k + (q & 1));
        const int tv[6][4] = {{v[0], v[7], v[3], v[1]}, {v[0], v[7], v[2], v[3]}, {v[0], v[7], v[6], v[2]},
                              {v[0], v[7], v[4], v[6]}, {v[0], v[7], v[5], v[4]}, {v[0], v[7], v[1], v[5]}};
        // periodic neighbour cubes
        int nb[7];
        nb[SAME] = c;
        nb[XM] = cube((i + cx - 1) % cx, j, k);
        nb[XP] = cube((i + 1) % cx, j, k);
        nb[YM] = cube(i, (j + cy - 1) % cy, k);
        nb[YP] = cube(i, (j + 1) % cy, k);
        nb[ZM] = cube(i, j, (k + cz - 1) % cz);
        nb[ZP] = cube(i, j, (k + 1) % cz);
        for (int l = 0; l < 6; l++) {
          int e = c * 6 + l;
          for (int q = 0; q < 4; q++) EtoV[e * 4 + q] = tv[l][q];
          for (int f = 0; f < 4; f++) {
            const Nbr &n = kNbr[l][f];
            EtoE[e * 4 + f] = nb[n.dir] * 6 + n.tet;
            EtoF[e * 4 + f] = n.face;
            EtoR[e * 4 + f] = n.rot;
            EtoB[e * 4 + f] = -99;
          }
        }
        // Non-periodic axes: the outer faces become boundary faces carrying
        // the physical tag (uniform_3d_single_core.cpp:217-272).
        auto mark = [&](int l, int f, int tag) {
          EtoE[(c * 6 + l) * 4 + f] = -99;
          EtoB[(c * 6 + l) * 4 + f] = tag;
        };
        if (phy_tags[0] >= 0) {
          if (i == 0) { mark(0, 3, phy_tags[0]); mark(1, 3, phy_tags[0]); }
          else if (i == cx - 1) { mark(3, 2, phy_tags[0]); mark(4, 2, phy_tags[0]); }
        }
        if (phy_tags[1] >= 0) {
          if (j == 0) { mark(4, 3, phy_tags[1]); mark(5, 3, phy_tags[1]); }
          else if (j == cy - 1) { mark(1, 2, phy_tags[1]); mark(2, 2, phy_tags[1]); }
        }
        if (phy_tags[2] >= 0) {
          if (k == 0) { mark(2, 3, phy_tags[2]); mark(3, 3, phy_tags[2]); }
          else if (k == cz - 1) { mark(0, 2, phy_tags[2]); mark(5, 2, phy_tags[2]); }
        }
      }
}

void element_points(int K, int Np, const int *EtoV, const double *VXYZ, const double *rst, double *PX, double *PY,
                    double *PZ) {
  for (int e = 0; e < K; e++) {
    const double *a = VXYZ + 3 * EtoV[e * 4 + 0], *b = VXYZ + 3 * EtoV[e * 4 + 1];
    const double *c = VXYZ + 3 * EtoV[e * 4 + 2], *d = VXYZ + 3 * EtoV[e * 4 + 3];
    for (int j = 0; j < Np; j++) {
      double r = rst[j * 3], s = rst[j * 3 + 1], t = rst[j * 3 + 2];
      double w0 = -(r + s + t + 1), w1 = r + 1, w2 = s + 1, w3 = t + 1;
      size_t o = (size_t)e * Np + j;
      PX[o] = 0.5 * (w0 * a[0] + w1 * b[0] + w2 * c[0] + w3 * d[0]);
      PY[o] = 0.5 * (w0 * a[1] + w1 * b[1] + w2 * c[1] + w3 * d[1]);
      PZ[o] = 0.5 * (w0 * a[2] + w1 * b[2] + w2 * c[2] + w3 * d[2]);
    }
  }
}

void element_metrics(int K, const int *EtoV, const double *VXYZ, double *J, double *const rst9[9]) {
  for (int e = 0; e < K; e++) {
    const double *a = VXYZ + 3 * EtoV[e * 4 + 0], *b = VXYZ + 3 * EtoV[e * 4 + 1];
    const double *c = VXYZ + 3 * EtoV[e * 4 + 2], *d = VXYZ + 3 * EtoV[e * 4 + 3];
    double xr = 0.5 * (b[0] - a[0]), xs = 0.5 * (c[0] - a[0]), xt = 0.5 * (d[0] - a[0]);
    double yr = 0.5 * (b[1] - a[1]), ys = 0.5 * (c[1] - a[1]), yt = 0.5 * (d[1] - a[1]);
    double zr = 0.5 * (b[2] - a[2]), zs = 0.5 * (c[2] - a[2]), zt = 0.5 * (d[2] - a[2]);
    double jac = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
    J[e] = jac;
    // inverse of the 3x3 map d(x,y,z)/d(r,s,t): rows r_x.., s_x.., t_x..
    rst9[0][e] = (ys * zt - zs * yt) / jac;
    rst9[1][e] = -(xs * zt - zs * xt) / jac;
    rst9[2][e] = (xs * yt - ys * xt) / jac;
    rst9[3][e] = -(yr * zt - zr * yt) / jac;
    rst9[4][e] = (xr * zt - zr * xt) / jac;
    rst9[5][e] = -(xr * yt - yr * xt) / jac;
    rst9[6][e] = (yr * zs - zr * ys) / jac;
    rst9[7][e] = -(xr * zs - zr * xs) / jac;
    rst9[8][e] = (xr * ys - yr * xs) / jac;
  }
}

void face_normals(int K, const double *J, const double *const m[9], double *nx, double *ny, double *nz, double *sJ,
                  double *Fscale) {
  for (int e = 0; e < K; e++) {
    double rx = m[0][e], ry = m[1][e], rz = m[2][e];
    double sx = m[3][e], sy = m[4][e], sz = m[5][e];
    double tx = m[6][e], ty = m[7][e], tz = m[8][e];
    // outward normals: face0 = -grad t, face1 = -grad s, face2 = grad(r+s+t), face3 = -grad r
    double n[4][3] = {{-tx, -ty, -tz}, {-sx, -sy, -sz}, {rx + sx + tx, ry + sy + ty, rz + sz + tz}, {-rx, -ry, -rz}};
    for (int f = 0; f < 4; f++) {
      size_t o = (size_t)e * 4 + f;
      double mag = std::sqrt(n[f][0] * n[f][0] + n[f][1] * n[f][1] + n[f][2] * n[f][2]);
      nx[o] = n[f][0] / mag;
      ny[o] = n[f][1] / mag;
      nz[o] = n[f][2] / mag;
      double sj = mag * J[e];
      sJ[o] = sj;
      Fscale[o] = sj / J[e];
    }
  }
}

void face_vmaps(int K, int Np, int Nfp, const int *EtoE, const int *EtoF, const int *EtoR, const int *Fmask,
                const int *FmaskP, int *vmapM, int *vmapP, int *vmapF) {
  for (int e = 0; e < K; e++)
    for (int f = 0; f < 4; f++) {
      int o = e * 4 + f;
      int nb = EtoE[o], nf = EtoF[o], rot = EtoR[o];
      for (int j = 0; j < Nfp; j++) vmapM[(size_t)o * Nfp + j] = e * Np + Fmask[f * Nfp + j];
      if (nb < 0) {
        vmapF[o] = o;
        for (int j = 0; j < Nfp; j++) vmapP[(size_t)o * Nfp + j] = vmapM[(size_t)o * Nfp + j];
      } else {
        vmapF[o] = nb * 4 + nf;
        for (int j = 0; j < Nfp; j++) vmapP[(size_t)o * Nfp + j] = nb * Np + FmaskP[(nf * 3 + rot) * Nfp + j];
      }
    }
}

namespace {
double tri_area(const double *p0, const double *p1, const double *p2) {
  double ux = p1[0] - p0[0], uy = p1[1] - p0[1], uz = p1[2] - p0[2];
  double vx = p2[0] - p0[0], vy = p2[1] - p0[1], vz = p2[2] - p0[2];
  double cx = uy * vz - uz * vy, cy = ux * vz - uz * vx, cz = ux * vy - uy * vx;
  return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}
}  // namespace

// Diameter of the inscribed sphere, 6V / (total face area) (mesh_3d_core.cpp:423-505).
void spatial_h(int K, const int *EtoV, const double *VXYZ, double *h) {
  for (int e = 0; e < K; e++) {
    const double *p0 = VXYZ + 3 * EtoV[e * 4 + 0], *p1 = VXYZ + 3 * EtoV[e * 4 + 1];
    const double *p2 = VXYZ + 3 * EtoV[e * 4 + 2], *p3 = VXYZ + 3 * EtoV[e * 4 + 3];
    double ax = p1[0] - p0[0], ay = p1[1] - p0[1], az = p1[2] - p0[2];
    double bx = p2[0] - p0[0], by = p2[1] - p0[1], bz = p2[2] - p0[2];
    double cx = p3[0] - p0[0], cy = p3[1] - p0[1], cz = p3[2] - p0[2];
    double vol = std::fabs(ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx)) / 6;
    double area = tri_area(p1, p0, p2) + tri_area(p0, p1, p3) + tri_area(p1, p2, p3) + tri_area(p2, p0, p3);
    h[e] = 6 * vol / area;
  }
}

}  // namespace dglm
