// Reference-tetrahedron operator tables (host side, extended precision).
#pragma once
#include <vector>

namespace dglm {

int tet_np(int N);  // (N+1)(N+2)(N+3)/6 volume nodes
int tri_np(int N);  // (N+1)(N+2)/2 nodes per face

// All tables are row-major doubles.  Shapes: rst/xyz [Np*3]; V, M, invM, Dr..Dt,
// WDr..WDt [Np*Np]; Emat, LIFT [Np*4Nfp]; Fmask [4*Nfp]; FmaskP [4*3*Nfp].
struct RefOps {
  int N = 0, Np = 0, Nfp = 0;
  std::vector<double> xyz, rst, V, M, invM, Dr, Ds, Dt, WDr, WDs, WDt, Emat, LIFT;
  std::vector<int> Fmask, FmaskP;
};

void reference_operators(int N, RefOps &out);
void interp_operators(int N, int high_N, std::vector<double> &high_rst, std::vector<double> &interpM,
                      std::vector<double> &high_M);

std::vector<long double> gll_nodes(int N);
long double jacobi_value(int n, long double a, long double b, long double x);
long double jacobi_grad(int n, long double a, long double b, long double x);
void tet_nodes_xyz(int N, std::vector<long double> &xyz);
void xyz_to_rst(const std::vector<long double> &xyz, std::vector<long double> &rst);
void rst_to_abc(const std::vector<long double> &rst, std::vector<long double> &abc);
void face_masks(int N, const std::vector<long double> &rst, int *Fmask, int *FmaskP);

}  // namespace dglm
