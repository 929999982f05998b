// Uniform tetrahedral mesh, affine geometry and face connectivity (host side).
#pragma once

namespace dglm {

void uniform_vertices(int nx, int ny, int nz, double x1, double x2, double y1, double y2, double z1, double z2,
                      double *VXYZ);
void uniform_connectivity(int nx, int ny, int nz, const int phy_tags[3], int *EtoV, int *EtoE, int *EtoF, int *EtoR,
                          int *EtoB);
void element_points(int K, int Np, const int *EtoV, const double *VXYZ, const double *rst, double *PX, double *PY,
                    double *PZ);
// metrics[9] = rx, ry, rz, sx, sy, sz, tx, ty, tz
void element_metrics(int K, const int *EtoV, const double *VXYZ, double *J, double *const metrics[9]);
void face_normals(int K, const double *J, const double *const metrics[9], double *nx, double *ny, double *nz,
                  double *sJ, double *Fscale);
void face_vmaps(int K, int Np, int Nfp, const int *EtoE, const int *EtoF, const int *EtoR, const int *Fmask,
                const int *FmaskP, int *vmapM, int *vmapP, int *vmapF);
void spatial_h(int K, const int *EtoV, const double *VXYZ, double *h);

}  // namespace dglm
