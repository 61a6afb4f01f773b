// Drives include/mdtk_gpu.hpp the way mdtk's SimLoop drives a potential:
// host-side atoms (the fields of mdtk::Atom), grad zeroed before each call,
// the velocity-Verlet eq.1-4 of SimLoop.cxx:323-399 on the host, and
// incDisplacement for every moved atom.  Reads a system from argv[1], runs
// argv[2] steps with dt = argv[3] s, writes energies per call and the final
// positions to argv[4].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mdtk_gpu.hpp"

struct V3 {
  double x = 0, y = 0, z = 0;
};
struct I3 {
  int x = 0, y = 0, z = 0;
};
struct Atom {
  int ID = 12;
  V3 coords, PBC, grad, V, an;
  I3 PBC_count;
  std::size_t globalIndex = 0;
  bool fixed = false;
  double M = 0;
  bool isFixed() const { return fixed; }
};
using AtomsArray = std::vector<Atom>;

static void wrap(double& c, double L, int& k) {
  if (L == 1e-4) return;
  while (c < 0) { c += L; --k; }
  while (c >= L) { c -= L; ++k; }
}

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  int n = 0;
  if (std::fread(&n, 4, 1, f) != 1) return 3;
  std::vector<int> elem(n);
  std::vector<double> x(3 * n), v(3 * n), pbc(3), mass(n);
  if (std::fread(elem.data(), 4, n, f) != (size_t)n || std::fread(x.data(), 8, 3 * n, f) != (size_t)(3 * n) ||
      std::fread(v.data(), 8, 3 * n, f) != (size_t)(3 * n) || std::fread(pbc.data(), 8, 3, f) != 3 ||
      std::fread(mass.data(), 8, n, f) != (size_t)n)
    return 3;
  std::fclose(f);
  int steps = std::atoi(argv[2]);
  double dt = std::atof(argv[3]);
  AtomsArray atoms(n);
  for (int i = 0; i < n; i++) {
    atoms[i].ID = elem[i];
    atoms[i].coords = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    atoms[i].V = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    atoms[i].PBC = {pbc[0], pbc[1], pbc[2]};
    atoms[i].globalIndex = i;
    atoms[i].M = mass[i];
  }
  std::vector<double> energies;
  try {
    mdtk_gpu::PotentialChain<AtomsArray> chain(MDG_POT_REBO | MDG_POT_AIREBO | MDG_POT_ETORS, true);
    auto force = [&]() {
      for (auto& a : atoms) a.grad = V3{};
      energies.push_back(chain(atoms));
    };
    force();
    for (auto& a : atoms) a.an = {-a.grad.x / a.M, -a.grad.y / a.M, -a.grad.z / a.M};
    for (int s = 0; s < steps; s++) {
      for (auto& a : atoms) {  // eq. 1
        V3 dr{a.V.x * dt + a.an.x * dt * dt / 2.0, a.V.y * dt + a.an.y * dt * dt / 2.0,
              a.V.z * dt + a.an.z * dt * dt / 2.0};
        a.coords.x += dr.x;
        a.coords.y += dr.y;
        a.coords.z += dr.z;
        chain.incDisplacement(a.globalIndex, dr.x, dr.y, dr.z);
      }
      force();
      for (auto& a : atoms) {  // eq. 2-4
        V3 h{a.V.x + a.an.x * dt / 2.0, a.V.y + a.an.y * dt / 2.0, a.V.z + a.an.z * dt / 2.0};
        a.an = {-a.grad.x / a.M, -a.grad.y / a.M, -a.grad.z / a.M};
        a.V = {h.x + a.an.x * dt / 2.0, h.y + a.an.y * dt / 2.0, h.z + a.an.z * dt / 2.0};
        wrap(a.coords.x, a.PBC.x, a.PBC_count.x);
        wrap(a.coords.y, a.PBC.y, a.PBC_count.y);
        wrap(a.coords.z, a.PBC.z, a.PBC_count.z);
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 4;
  }
  FILE* o = std::fopen(argv[4], "wb");
  int ne = (int)energies.size();
  std::fwrite(&ne, 4, 1, o);
  std::fwrite(energies.data(), 8, ne, o);
  for (auto& a : atoms) std::fwrite(&a.coords, 8, 3, o);
  std::fclose(o);
  std::printf("adapter_check: %d force calls, E0 = %.12e erg\n", ne, energies[0]);
  return 0;
}
