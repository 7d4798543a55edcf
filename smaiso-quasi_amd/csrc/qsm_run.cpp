// qsm_run -- host driver over libqsm: the counterpart of the reference's
// main() (quasi/quasi.cpp:98-2325, quasi/quasi-stb.cpp:93-2215) for one GPU.
//
//   qsm_run [params-file] [key=value ...]
//
// Keys are the qsm_params fields (nx ny nz variant compat_ranks nw divv_switch
// hydro q0 alpha beta gamma eps nu kp rho0 amp rho_m dt filter growth_q) and:
//   prec=fp64|mixed        arithmetic (default fp64)
//   ic=interface|tilted|stb|focal|load|ckpt
//                          initial condition (quasi.cpp:504-589, stb:529-552);
//                          load reads psi<rank>.dat text files of a P-rank
//                          reference run (quasi.cpp:692-717); ckpt resumes a
//                          binary checkpoint
//   load=DIR ranks=P       where/how many psi<rank>.dat files (ic=load)
//   ckpt_in=PATH ckpt_out=PATH
//   steps=N                stop after N steps (0: run until L1 <= lim_l1)
//   step_l1=50 step_save=10 lim_l1=1e-6   diagnostics cadence (quasi.cpp:159-160, 969)
//   out=DIR                output directory (reference file names and formats:
//                          info, L1.dat, mass.dat, energy.dat, psiMid.dat,
//                          vx.dat, vy.dat, vz.dat, divv.dat, p.dat, rho.dat,
//                          psi<rank>.dat; stb: L1.dat, mass.dat, growth.dat,
//                          psiMid.dat at 8 decimals, dtpsi.dat, fourier.dat)
//   fc=SIZE                focal-conic size aE = bE (argv[3] of the reference)
//   noise=1e-4 seed=1      Gaussian noise added to psi (std::mt19937_64)
//   graphs=1               replay steps as HIP graphs
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qsm.h"

namespace {

std::map<std::string, std::string> g_kv;

void parse_line(const std::string& line) {
  std::string s = line.substr(0, line.find('#'));
  const size_t eq = s.find('=');
  if (eq == std::string::npos) return;
  auto trim = [](std::string t) {
    const size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
  };
  g_kv[trim(s.substr(0, eq))] = trim(s.substr(eq + 1));
}

std::string get(const std::string& k, const std::string& def) {
  auto it = g_kv.find(k);
  return it == g_kv.end() ? def : it->second;
}
double getd(const std::string& k, double def) {
  auto it = g_kv.find(k);
  return it == g_kv.end() ? def : std::atof(it->second.c_str());
}
long getl(const std::string& k, long def) {
  auto it = g_kv.find(k);
  return it == g_kv.end() ? def : std::atol(it->second.c_str());
}

int check(int rc, const char* what) {
  if (rc != QSM_OK) {
    std::fprintf(stderr, "qsm_run: %s failed (%d): %s\n", what, rc, qsm_last_error());
    std::exit(1);
  }
  return rc;
}

size_t at(const qsm_params& p, int i, int j, int k) { return ((size_t)i * p.ny + j) * p.nz + k; }

// ---- initial conditions (host, FP64)

void ic_interface(const qsm_params& p, std::vector<double>& psi) {
  const double dz = 2 * M_PI / (p.q0 * p.nw), lz = p.nz * dz;
  for (int i = 0; i < p.nx; ++i)
    for (int j = 0; j < p.ny; ++j)
      for (int k = 0; k < p.nz; ++k) {
        const double z = k * dz;
        psi[at(p, i, j, k)] = p.amp * std::cos(p.q0 * z) * 0.5 * (1.0 - std::tanh((z - lz / 2) / 2.0));
      }
}

void ic_tilted(const qsm_params& p, std::vector<double>& psi) {
  const double d = 2 * M_PI / (p.q0 * p.nw), t = 15.0 * M_PI / 180.0;
  const double lx = p.nx * d, ly = p.ny * d, lz = p.nz * d;
  const double u0 = (lz / 2) * std::cos(t) + (ly / 2) * std::sin(t);
  for (int i = 0; i < p.nx; ++i) {
    const double shift = 2 * d * std::sin(2 * M_PI * (i * d) / (lx / 4));
    for (int j = 0; j < p.ny; ++j)
      for (int k = 0; k < p.nz; ++k) {
        const double u = k * d * std::cos(t) + j * d * std::sin(t);
        psi[at(p, i, j, k)] =
            p.amp * std::cos(p.q0 * u) * 0.5 * (1.0 - std::tanh((u - u0 - shift) / 2.0));
      }
  }
}

// quasi-stb.cpp:546-547 with phi = 0.1
void ic_stb(const qsm_params& p, std::vector<double>& psi, double phi) {
  const double dx = 2 * M_PI / (p.q0 * p.nw);
  for (int i = 0; i < p.nx; ++i)
    for (int j = 0; j < p.ny; ++j)
      for (int k = 0; k < p.nz; ++k)
        psi[at(p, i, j, k)] = p.amp * std::cos(p.q0 * k * dx) +
                              (p.amp * phi / 2) * std::cos(p.q0 * k * dx + p.growth_q * i * dx) -
                              (p.amp * phi / 2) * std::cos(p.q0 * k * dx - p.growth_q * i * dx);
}

// Focal-conic domain of quasi.cpp:543-589 (aE = bE = fc; z measured from the
// bottom; layers A cos(q0 dz r') inside the torus-like region, flat layers
// outside it up to z < bE, isotropic above).
void ic_focal(const qsm_params& p, std::vector<double>& psi, double fc) {
  const double dz = 2 * M_PI / (p.q0 * p.nw), mid = p.nx / 2, aE = fc, bE = fc;
  for (int i = 0; i < p.nx; ++i)
    for (int j = 0; j < p.ny; ++j)
      for (int k = 0; k < p.nz; ++k) {
        double v = 0.0;
        if (k < bE + 1) {
          const double xs = i - mid, ys = j - mid, zs = k;
          const double ds = std::sqrt(xs * xs + ys * ys);
          if (ds < mid) {
            if (std::sqrt(std::pow((ds - mid) / aE, 2) + std::pow(zs / bE, 2)) <= 1)
              v = p.amp * std::cos(p.q0 * dz * std::sqrt(std::pow((bE / aE) * (ds - mid), 2) + zs * zs));
          } else if (std::abs(zs) < bE) {
            v = p.amp * std::cos(p.q0 * zs * dz);
          }
        }
        psi[at(p, i, j, k)] = v;
      }
}

// psi<rank>.dat of a P-rank reference run: one value per line, [i_local][j][k]
void ic_load(const qsm_params& p, std::vector<double>& psi, const std::string& dir, int ranks) {
  const int n0 = p.nx / ranks;
  for (int r = 0; r < ranks; ++r) {
    const std::string path = dir + "/psi" + std::to_string(r) + ".dat";
    std::ifstream in(path);
    if (!in) { std::fprintf(stderr, "qsm_run: cannot read %s\n", path.c_str()); std::exit(1); }
    for (int il = 0; il < n0; ++il)
      for (int j = 0; j < p.ny; ++j)
        for (int k = 0; k < p.nz; ++k)
          if (!(in >> psi[at(p, r * n0 + il, j, k)])) {
            std::fprintf(stderr, "qsm_run: %s is short\n", path.c_str());
            std::exit(1);
          }
  }
}

// ---- reference-format outputs

void append_value(const std::string& path, double v) {
  std::ofstream o(path, std::ios_base::app);
  o << v << "\n";
}

// mid-plane j = Ny/2 slice in [i][k] order, appended (quasi.cpp:2040-2273,
// quasi-stb.cpp:2040-2108); fixed8: std::fixed with 8 decimals, the stb
// psiMid.dat format (quasi-stb.cpp:2065), else the stream default (6
// significant digits)
void append_mid(qsm_handle h, const qsm_params& p, int which, const std::string& path,
                std::vector<double>& buf, bool fixed8 = false) {
  check(qsm_get_field(h, which, buf.data()), "get_field");
  std::ofstream o(path, std::ios_base::app);
  if (fixed8) o << std::fixed << std::setprecision(8);
  const int j = p.ny / 2;
  for (int i = 0; i < p.nx; ++i)
    for (int k = 0; k < p.nz; ++k) o << buf[at(p, i, j, k)] << "\n";
}

// full psi per reference rank, overwritten (quasi.cpp:2051-2062)
void write_psi_ranks(const qsm_params& p, const std::vector<double>& psi, const std::string& dir,
                     int ranks) {
  const int n0 = p.nx / ranks;
  for (int r = 0; r < ranks; ++r) {
    std::ofstream o(dir + "/psi" + std::to_string(r) + ".dat");
    for (int il = 0; il < n0; ++il)
      for (int j = 0; j < p.ny; ++j)
        for (int k = 0; k < p.nz; ++k) o << psi[at(p, r * n0 + il, j, k)] << "\n";
  }
}

}  // namespace

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    if (std::strchr(argv[a], '=')) {
      parse_line(argv[a]);
    } else {
      std::ifstream in(argv[a]);
      if (!in) { std::fprintf(stderr, "qsm_run: cannot read %s\n", argv[a]); return 1; }
      std::string line;
      while (std::getline(in, line)) parse_line(line);
    }
  }
  for (int a = 1; a < argc; ++a)  // command-line keys override the file
    if (std::strchr(argv[a], '=')) parse_line(argv[a]);

  qsm_params p;
  const int variant = get("variant", "0") == "1" || get("variant", "") == "stb" ? QSM_STB : QSM_QUASI;
  check(qsm_default_params(variant, &p), "default_params");
  p.nx = (int)getl("nx", p.nx); p.ny = (int)getl("ny", p.ny); p.nz = (int)getl("nz", p.nz);
  p.compat_ranks = (int)getl("compat_ranks", p.compat_ranks);
  p.nw = (int)getl("nw", p.nw); p.divv_switch = (int)getl("divv_switch", p.divv_switch);
  p.hydro = (int)getl("hydro", p.hydro);
  p.q0 = getd("q0", p.q0); p.alpha = getd("alpha", p.alpha); p.beta = getd("beta", p.beta);
  p.gamma = getd("gamma", p.gamma); p.eps = getd("eps", p.eps); p.nu = getd("nu", p.nu);
  p.kp = getd("kp", p.kp); p.rho0 = getd("rho0", p.rho0); p.amp = getd("amp", p.amp);
  p.rho_m = getd("rho_m", p.rho_m); p.dt = getd("dt", p.dt); p.filter = getd("filter", p.filter);
  p.growth_q = getd("growth_q", p.growth_q);
  const int prec = get("prec", "fp64") == "mixed" ? QSM_MIXED : QSM_FP64;
  const long steps = getl("steps", 0);
  const int step_l1 = (int)getl("step_l1", variant == QSM_STB ? 1 : 50);
  const int step_save = (int)getl("step_save", variant == QSM_STB ? 1 : 10);
  const double lim_l1 = getd("lim_l1", 1e-6);
  const std::string out = get("out", "");
  const int ranks = (int)getl("ranks", p.compat_ranks > 0 ? p.compat_ranks : 1);

  qsm_handle h;
  check(qsm_create(&p, prec, &h), "create");
  qsm_set_graphs(h, (int)getl("graphs", 1));
  const size_t n = (size_t)p.nx * p.ny * p.nz;
  std::vector<double> psi(n, 0.0), buf(n);

  const std::string ic = get("ic", variant == QSM_STB ? "stb" : "interface");
  if (ic == "ckpt") {
    check(qsm_checkpoint_read(h, get("ckpt_in", "").c_str()), "checkpoint_read");
  } else {
    if (ic == "interface") ic_interface(p, psi);
    else if (ic == "tilted") ic_tilted(p, psi);
    else if (ic == "stb") ic_stb(p, psi, getd("phi", 0.1));
    else if (ic == "focal") ic_focal(p, psi, getd("fc", 32));
    else if (ic == "load") ic_load(p, psi, get("load", "."), ranks);
    else { std::fprintf(stderr, "qsm_run: unknown ic %s\n", ic.c_str()); return 1; }
    const double noise = getd("noise", ic == "interface" || ic == "tilted" ? 1e-4 : 0.0);
    if (noise > 0) {
      std::mt19937_64 rng((unsigned long long)getl("seed", 1));
      std::normal_distribution<double> nd(0.0, noise);
      for (auto& v : psi) v += nd(rng);
    }
    check(qsm_set_psi(h, psi.data()), "set_psi");
  }

  if (!out.empty()) {
    mkdir(out.c_str(), 0755);
    std::ofstream info(out + "/info");  // quasi.cpp:604-631
    info << "Nx: " << p.nx << "\n" << "Ny: " << p.ny << "\n" << "Nz: " << p.nz << "\n"
         << "Points per wavelength (nw): " << p.nw << "\n" << "q0 (z direction): " << p.q0 << "\n"
         << "epsilon: " << p.eps << "\n" << "beta: " << p.beta << "\n" << "gamma: " << p.gamma << "\n"
         << "Viscosity (nu): " << p.nu << "\n" << "dx: " << 2 * M_PI / (p.q0 * p.nw) << "\n"
         << "dt: " << p.dt << "\n" << "stepL1: " << step_l1 << "\n" << "stepSave: " << step_save << "\n"
         << "divv switch: " << p.divv_switch << "\n" << "Initial amp: " << p.amp << "\n"
         << "kp: " << p.kp << "\n" << "rho_0: " << p.rho0 << "\n"
         << "Precision: " << (prec == QSM_MIXED ? "mixed" : "fp64") << "\n";
    // the files a run starts empty (quasi.cpp:633-687; quasi-stb.cpp:646-690)
    const std::vector<const char*> fresh =
        variant == QSM_QUASI
            ? std::vector<const char*>{"L1.dat", "mass.dat", "energy.dat", "psiMid.dat", "vx.dat",
                                       "vy.dat", "vz.dat", "divv.dat", "rho.dat", "p.dat"}
            : std::vector<const char*>{"growth.dat", "L1.dat", "mass.dat", "psiMid.dat",
                                       "dtpsi.dat", "fourier.dat"};
    for (const char* f : fresh) std::ofstream(out + "/" + f);
  }

  const auto t0 = std::chrono::steady_clock::now();
  long done = 0;
  int saves = 0;
  double L1 = 1.0;  // quasi.cpp:963
  while ((steps > 0 && done < steps) || (steps == 0 && L1 > lim_l1)) {
    const long k = steps > 0 ? std::min<long>(step_l1, steps - done) : step_l1;
    check(qsm_step(h, (int)k), "step");
    done += k;
    qsm_diag d;
    const int rc = qsm_get_diag(h, &d);
    if (rc == QSM_ERR_NONFINITE) {
      std::fprintf(stderr, "qsm_run: the run blew up at step %ld: %s\n", done, qsm_last_error());
      return 2;
    }
    check(rc, "get_diag");
    L1 = d.L1;
    if (!out.empty()) {
      append_value(out + "/L1.dat", d.L1);
      append_value(out + "/mass.dat", d.mass);
      if (variant == QSM_QUASI) append_value(out + "/energy.dat", d.energy);
      if (variant == QSM_STB && done > 1) append_value(out + "/growth.dat", d.growth);
      if (++saves == step_save) {
        saves = 0;
        check(qsm_get_field(h, QSM_PSI, psi.data()), "get_field");
        write_psi_ranks(p, psi, out, ranks);
        if (variant == QSM_QUASI) {  // quasi.cpp:2040-2273
          const std::pair<int, const char*> mids[] = {
              {QSM_PSI, "psiMid.dat"}, {QSM_VX, "vx.dat"}, {QSM_VY, "vy.dat"}, {QSM_VZ, "vz.dat"},
              {QSM_DIVV, "divv.dat"}, {QSM_P, "p.dat"}, {QSM_RHO, "rho.dat"}};
          for (const auto& m : mids) append_mid(h, p, m.first, out + "/" + m.second, buf);
        } else {  // quasi-stb.cpp:2040-2108 (its v slices are commented out)
          append_mid(h, p, QSM_PSI, out + "/psiMid.dat", buf, true);
          append_mid(h, p, QSM_DTPSI, out + "/dtpsi.dat", buf);
        }
        if (variant == QSM_STB) {  // spectral slice j = 0, 16 digits (quasi-stb.cpp:2127-2139)
          check(qsm_get_field(h, QSM_PSIHAT, buf.data()), "get_field");
          std::ofstream o(out + "/fourier.dat", std::ios_base::app);
          for (int i = 0; i < p.nx; ++i)
            for (int kk = 0; kk < p.nz; ++kk)
              o << std::fixed << std::setprecision(16) << buf[at(p, i, 0, kk)] << "\n";
        }
      }
    }
    std::printf("step %ld L1 %.6e mass %.10e energy %.10e\n", done, d.L1, d.mass, d.energy);
  }
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("steps %ld wall %.3f s  %.3e grid-point updates/s\n", done, wall,
              (double)n * done / (wall > 0 ? wall : 1));
  const std::string ck = get("ckpt_out", "");
  if (!ck.empty()) check(qsm_checkpoint_write(h, ck.c_str()), "checkpoint_write");
  qsm_destroy(h);
  return 0;
}
