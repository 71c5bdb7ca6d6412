// Host code under AddressSanitizer and UBSan (built and run by
// tests/test_host_sanitizers.py): mesh build at levels 0-5, netCDF mesh
// files written and read back, every truncation of a file and 3000 random
// byte corruptions of its header (rejected or read, never a memory error),
// VTK mesh output in the four layouts, and the initial-condition routines.
// Files go to the directory given as argv[1].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>
#include "lpm_bve.h"

static std::vector<char> slurp(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
static void spit(const std::string& p, const std::vector<char>& b, size_t n) {
  std::ofstream f(p, std::ios::binary | std::ios::trunc);
  f.write(b.data(), (std::streamsize)n);
}
static int read_all(const char* path) {
  int64_t c[4];
  if (lpm_mesh_read_nc(path, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c)) return 1;
  if (c[0] < 0 || c[2] < 1 || c[0] > (1 << 24) || c[2] > (1 << 24)) return 1;
  std::vector<double> vx(3 * c[0]), vl(3 * c[0]), fx(3 * c[2]), fl(3 * c[2]), fa(c[2]);
  std::vector<uint8_t> fm(c[2]);
  std::vector<int32_t> fv(3 * c[2]), fp(c[2]);
  return lpm_mesh_read_nc(path, vx.data(), vl.data(), fx.data(), fl.data(), fa.data(), fm.data(), fv.data(),
                          fp.data(), c);
}
int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  const std::string mesh = dir + "/mesh.nc", cut = dir + "/cut.nc", fuzz = dir + "/fuzz.nc", vtk = dir + "/mesh.vtk";
  for (int lev = 0; lev <= 5; ++lev) {
    int64_t nv, ne, nf;
    lpm_icos_alloc(lev, &nv, &ne, &nf);
    std::vector<double> vx(3 * nv), vl(3 * nv), fx(3 * nf), fl(3 * nf), fa(nf);
    std::vector<uint8_t> fm(nf);
    std::vector<int32_t> fv(3 * nf), fp(nf);
    int64_t counts[4];
    if (lpm_icos_tree_init(lev, nullptr, vx.data(), vl.data(), fx.data(), fl.data(), fa.data(), fm.data(), fv.data(),
                           fp.data(), counts)) return 2;
  }
  const char* p = mesh.c_str();
  for (int lev = 0; lev <= 3; ++lev) {
    if (lpm_mesh_write_nc(lev, nullptr, p)) return 3;
    if (read_all(p)) return 4;
    for (int fmt = 0; fmt < 4; ++fmt)
      if (lpm_mesh_write_vtk(lev, nullptr, vtk.c_str(), fmt)) return 5;
  }
  std::vector<char> b = slurp(p);
  int ok = 0, bad = 0;
  for (size_t n = 0; n < b.size(); n += (n < 2048 ? 1 : 97)) {
    spit(cut, b, n);
    (read_all(cut.c_str()) ? bad : ok)++;
  }
  std::mt19937 rng(7);
  for (int k = 0; k < 3000; ++k) {
    std::vector<char> c = b;
    const int nflip = 1 + (int)(rng() % 4);
    for (int j = 0; j < nflip; ++j) c[rng() % std::min<size_t>(c.size(), 1500)] = (char)(rng() & 0xFF);
    spit(fuzz, c, c.size());
    (read_all(fuzz.c_str()) ? bad : ok)++;
  }
  std::vector<double> xyz = {0, 0, 1, 1, 0, 0, 0.6, 0.8, 0}, z(3), f(3), u(9), un(4), nn(4);
  double cen[6] = {0, 0, 1, 1, 0, 0}, amps[2] = {1, -1};
  if (lpm_ic_eval(xyz.data(), 3, z.data(), 0.5, 4, 0.1, 0.1, 2, cen, amps, 0.2)) return 6;
  if (lpm_poisson_y54(xyz.data(), 3, f.data(), u.data())) return 7;
  if (lpm_rng_draws(3, 4, un.data(), 4, nn.data())) return 8;
  std::printf("sanitized host run ok: %d files read, %d rejected\n", ok, bad);
  return 0;
}
