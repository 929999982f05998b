// The datain/ file contract (host side): reading a work directory written by
// the launcher's save_to_binaries (dim3/src/python/dglm_3d_launcher.py:74-149;
// the reference's own reader is read_bin_file, utils/src/cpp/io.cpp:7-17,
// called from DataHost, euler_data.cpp:71-178), and the face connectivity
// proven from its vmapP / vmapF arrays.
#pragma once
#include <string>
#include <vector>

namespace dglm {

// Status codes as the C ABI's DGLM_ERR_* values.
enum { HOST_OK = 0, HOST_ERR_ARG = -1, HOST_ERR_IO = -3, HOST_ERR_UNSUPPORTED = -5 };

struct Datain {
  int K = 0;
  int real_bytes = 0;  // 8 ('f8' files) or 4 ('f4' files, widened exactly)
  std::vector<int> EtoB, vmapP, vmapF, Fmask;
  std::vector<double> WDr, WDs, WDt, LIFT, Fscale, metric[9], nx, ny, nz, px, py, pz;
};

// Reads dir/{dim,EtoB,vmapP,vmapF,Fmask,WDr,WDs,WDt,LIFT,Fscale,rx..tz,nx,ny,
// nz,px,py,pz}.bin for order N (Np, Nfp the order's node counts).  Integer
// files are int32; every floating-point file is f8, or every one f4.  A
// missing, short, long or mixed-width file is HOST_ERR_IO with a message.
int read_datain(const std::string &dir, int Np, int Nfp, Datain &d, std::string &err);

// Packed face connectivity (nbr << 4) | (nbr_face << 2) | code from vmapF and
// vmapP: code 0..2 is the rotation r with vmapP[e,f,:] = nbr*Np +
// fmx[(nbr_face*4 + r)*Nfp + :] (FmaskP order), code 3 the identity (Fmask
// order, a face that is its own neighbour).  fmx [16*Nfp] holds FmaskP[f][r]
// at code r < 3 and Fmask[f] at code 3.  A vmapF out of range is
// HOST_ERR_ARG; a vmapP that is no rotation of its neighbour face is
// HOST_ERR_UNSUPPORTED.  EtoB (optional): faces tagged 1 must be their own
// neighbour (Dirichlet faces); dmap[K*4] gets their slot (-1 elsewhere) and
// dface their face indices in ascending order.
int pack_connectivity(int K, int Np, int Nfp, const int *vmapP, const int *vmapF, const int *EtoB, const int *fmx,
                      std::vector<int> &conn, std::vector<int> &dmap, std::vector<int> &dface, std::string &err);

}  // namespace dglm
