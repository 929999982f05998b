// Driver around the reference CPU solver (TEST INFRASTRUCTURE ONLY).
//
// Compiled together with the reference's own advec_* / euler_* sources via
// the generated common.h (see Makefile).  Modes:
//   run                      the reference main loop (euler_main.cpp:3-17):
//                            init, runRK until TMAX, print "tstep=.., dt=.., t=.."
//                            and write dataout/u*_NNNNNNN.bin on the cadence
//   steps <n> <u0|-> <out>   fixed step count (the main10 pattern,
//                            euler_nopad_main10.cu:4-17); u0 '-' means the
//                            reference initial condition; writes all
//                            variables, the dt history and t to <out>
//   rhs <u> <out>            one RHS at state u: calc_maxv, update_tau,
//                            update_ub, calc_fluxLM, calc_fluxes, update_rhs;
//                            writes utmp, maxv and tau to <out>
// Environment: DGLM_DATAIN, DGLM_DATAOUT, DGLM_TMAX, DGLM_CFL, DGLM_GAMMA,
// DGLM_TAU_SCALE, and for advection DGLM_VEL (file of vx,vy,vz [3*K*Np]).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

std::string g_datain, g_dataout;
double g_tmax, g_cfl, g_gamma, g_tau_scale;

static std::string env_s(const char *k, const char *d) {
  const char *v = getenv(k);
  return v ? std::string(v) : std::string(d);
}
static double env_d(const char *k, double d) {
  const char *v = getenv(k);
  return v ? atof(v) : d;
}

#include "common.h"

static bool read_raw(const std::string &path, double *dst, size_t n) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  size_t got = fread(dst, sizeof(double), n, f);
  fclose(f);
  return got == n;
}

static double *state_ptr(DataHost &host, int v) {
#if NVAR == 5
  return host.p_u[v].data();
#else
  (void)v;
  return host.u.data();
#endif
}

static double *utmp_ptr(DataHost &host, int v) {
#if NVAR == 5
  return host.p_utmp[v].data();
#else
  (void)v;
  return host.utmp.data();
#endif
}

static void load_velocity(DataHost &host) {
#if NVAR == 1
  const size_t n = (size_t)host.nelem * NP;
  for (size_t i = 0; i < n; i++) {
    host.vx[i] = 1.;
    host.vy[i] = 1.;
    host.vz[i] = 1.;
  }
  std::string vf = env_s("DGLM_VEL", "");
  if (!vf.empty()) {
    std::vector<double> v(3 * n);
    if (!read_raw(vf, v.data(), 3 * n)) {
      fprintf(stderr, "cannot read %s\n", vf.c_str());
      exit(2);
    }
    for (size_t i = 0; i < n; i++) {
      host.vx[i] = v[i];
      host.vy[i] = v[n + i];
      host.vz[i] = v[2 * n + i];
    }
  }
#else
  (void)host;
#endif
}

static void load_state(DataHost &host, const std::string &path) {
  const size_t n = (size_t)host.nelem * NP;
  std::vector<double> u((size_t)NVAR * n);
  if (!read_raw(path, u.data(), u.size())) {
    fprintf(stderr, "cannot read %s\n", path.c_str());
    exit(2);
  }
  for (int v = 0; v < NVAR; v++)
    for (size_t i = 0; i < n; i++) state_ptr(host, v)[i] = u[v * n + i];
}

static void one_rhs(DataHost &host) {
#if NVAR == 5
  calc_maxv(host.nelem, host.Fmask.data(), host.Fscale.data(), host.p_u.data(), host.maxv.data(),
            host.maxv_fs.data());
  update_tau(host.nelem, host.vmapF.data(), host.maxv.data(), host.tau.data());
  update_ub(host.nelem, host.Fmask.data(), host.vmapP.data(), host.p_u.data(), host.p_ub.data());
  calc_fluxLM(host.nelem, 0.0, host.Fmask.data(), host.EtoB.data(), host.nx.data(), host.ny.data(), host.nz.data(),
              host.px.data(), host.py.data(), host.pz.data(), host.tau.data(), host.Fscale.data(), host.p_ub.data(),
              host.p_u.data(), host.p_fluxLM.data());
  calc_fluxes(host.nelem, host.rx.data(), host.ry.data(), host.rz.data(), host.sx.data(), host.sy.data(),
              host.sz.data(), host.tx.data(), host.ty.data(), host.tz.data(), host.p_u.data(), host.p_fluxR.data(),
              host.p_fluxS.data(), host.p_fluxT.data());
  update_rhs(host.nelem, host.WDr.data(), host.WDs.data(), host.WDt.data(), host.LIFT.data(), host.p_fluxR.data(),
             host.p_fluxS.data(), host.p_fluxT.data(), host.p_fluxLM.data(), host.p_utmp.data());
#else
  calc_maxv(host.nelem, host.Fmask.data(), host.Fscale.data(), host.vx.data(), host.vy.data(), host.vz.data(),
            host.maxv.data(), host.maxv_fs.data());
  update_tau(host.nelem, host.vmapF.data(), host.maxv.data(), host.tau.data());
  update_ub(host.nelem, host.Fmask.data(), host.vmapP.data(), host.u.data(), host.ub.data());
  calc_fluxLM(host.nelem, 0.0, host.Fmask.data(), host.EtoB.data(), host.nx.data(), host.ny.data(), host.nz.data(),
              host.px.data(), host.py.data(), host.pz.data(), host.vx.data(), host.vy.data(), host.vz.data(),
              host.tau.data(), host.Fscale.data(), host.ub.data(), host.u.data(), host.fluxLM.data());
  calc_fluxes(host.nelem, host.rx.data(), host.ry.data(), host.rz.data(), host.sx.data(), host.sy.data(),
              host.sz.data(), host.tx.data(), host.ty.data(), host.tz.data(), host.vx.data(), host.vy.data(),
              host.vz.data(), host.u.data(), host.fluxR.data(), host.fluxS.data(), host.fluxT.data());
  update_rhs(host.nelem, host.WDr.data(), host.WDs.data(), host.WDt.data(), host.LIFT.data(), host.fluxR.data(),
             host.fluxS.data(), host.fluxT.data(), host.fluxLM.data(), host.utmp.data());
#endif
}

int main(int argc, char **argv) {
  g_datain = env_s("DGLM_DATAIN", "datain/");
  g_dataout = env_s("DGLM_DATAOUT", "dataout/");
  g_tmax = env_d("DGLM_TMAX", 1000.0);
  g_cfl = env_d("DGLM_CFL", 0.8);
  g_gamma = env_d("DGLM_GAMMA", 1.4);
  g_tau_scale = env_d("DGLM_TAU_SCALE", 1.2);
  if (argc < 2) {
    fprintf(stderr, "usage: %s run | steps n u0|- out | rhs u out\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  DataHost host;
  const size_t n = (size_t)host.nelem * NP;

  if (mode == "run") {
    init(host);
    load_velocity(host);  // velocity (1,1,1) unless DGLM_VEL is given
    while (host.t < TMAX) {
      runRK(host);
      if (host.tstep % host.print_tstep == 0 or host.last_tstep) {
        printf("tstep=%d, dt=%.8f, t=%.5f\n", host.tstep, host.dt, host.t);
        host.write_u(host.tstep);
      }
    }
    return 0;
  }

  if (mode == "steps" && argc >= 5) {
    int nsteps = atoi(argv[2]);
    std::string u0 = argv[3];
    if (u0 == "-") init(host);
    load_velocity(host);
    if (u0 != "-") load_state(host, u0);
    std::vector<double> dts;
    for (int s = 0; s < nsteps && host.t < TMAX; s++) {
      runRK(host);
      dts.push_back(host.dt);
    }
    FILE *f = fopen(argv[4], "wb");
    if (!f) return 3;
    for (int v = 0; v < NVAR; v++) fwrite(state_ptr(host, v), sizeof(double), n, f);
    double t = host.t, cnt = (double)dts.size();
    fwrite(&t, sizeof(double), 1, f);
    fwrite(&cnt, sizeof(double), 1, f);
    fwrite(dts.data(), sizeof(double), dts.size(), f);
    fclose(f);
    return 0;
  }

  if (mode == "rhs" && argc >= 4) {
    load_velocity(host);
    load_state(host, argv[2]);
    one_rhs(host);
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 3;
    for (int v = 0; v < NVAR; v++) fwrite(utmp_ptr(host, v), sizeof(double), n, f);
    fwrite(host.maxv.data(), sizeof(double), host.maxv.size(), f);
    fwrite(host.tau.data(), sizeof(double), host.tau.size(), f);
    fclose(f);
    return 0;
  }
  fprintf(stderr, "bad arguments\n");
  return 2;
}
