"""Write the common.h that composes the reference CPU solver (test infrastructure).

The reference generates this header per problem (dglm_3d_launcher.py:152-205):
#defines for the problem constants followed by #includes of io.cpp and the
chosen {physics}_{data,dglm,time,init}.cpp files.  Here the run-time knobs
(datain/dataout directories, TMAX, CFL, GAMMA, TAU_SCALE) are bound to globals
that ref_driver.cpp fills from the environment, so one executable per
physics and order serves every test case.

usage: gen_ref_common.py {advec|euler}_N<n> <out-path>
"""
import sys


def main():
    tag, out = sys.argv[1], sys.argv[2]
    phys, n = tag.split("_N")
    N = int(n)
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    nvar = 5 if phys == "euler" else 1
    lines = [
        "#ifndef COMMON_H", "#define COMMON_H", "",
        "#include <string>",
        "extern std::string g_datain, g_dataout;",
        "extern double g_tmax, g_cfl, g_gamma, g_tau_scale;", "",
        "#define REAL    double",
        "#define DATAIN  g_datain",
        "#define DATAOUT g_dataout", "",
        "#define PI     3.14159265358979323846",
        "#define TDIGIT 7",
        f"#define N      {N}",
        f"#define NP     {Np}",
        f"#define NFP    {Nfp}",
        "#define NFACE  4",
        "#define NROT   3",
        f"#define NVAR   {nvar}",
        "#define TMAX   g_tmax",
        "#define CFL    g_cfl",
        "#define GAMMA  g_gamma",
        "#define TAU_SCALE g_tau_scale", "",
        '#include "io.cpp"',
    ] + [f'#include "{phys}_{part}.cpp"' for part in ("data", "dglm", "time", "init")] + ["", "#endif", ""]
    with open(out, "w") as f:
        f.write("\n".join(lines))


if __name__ == "__main__":
    main()
