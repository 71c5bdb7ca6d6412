// The reference's SpherePoisson test program (tests/LpmSpherePoisson.cpp:9-55)
// on this library: for each tree depth, on the icosahedral and then the
// cubed-sphere seed, build the mesh, set the Y_5^4 right-hand side, solve on
// the GPU and print the mesh size and L-infinity errors in the reference's
// format.  With a file root, each solution is also written as
// SpherePoisson::outputVtk writes it (examples/SphereVorticity.cpp:62-65):
// <root>icostri<depth>.vtk and <root>cubedsphere<depth>.vtk.
//
//     poisson_driver [max_depth] [vtk_root]     (depths 0 .. max_depth, default 3)
#include <cstdlib>
#include <iostream>
#include <string>

#include "lpm_bve.hpp"

using namespace Lpm;

int main(int argc, char* argv[]) {
  const int max_depth = argc > 1 ? std::atoi(argv[1]) : 3;
  const std::string vtk_root = argc > 2 ? argv[2] : "";
  try {
    for (int tree_depth = 0; tree_depth <= max_depth; ++tree_depth) {
      Index nmaxverts, nmaxedges, nmaxfaces;
      {
        MeshSeed<IcosTriSphereSeed> triseed;
        triseed.setMaxAllocations(nmaxverts, nmaxedges, nmaxfaces, tree_depth);
        SpherePoisson<IcosTriSphereSeed> ic(nmaxverts, nmaxedges, nmaxfaces);
        ic.treeInit(tree_depth, triseed);
        ic.updateDevice();
        ic.init();
        std::cout << "icostri: ";
        ic.solve();
        ic.updateHost();
        if (!vtk_root.empty()) ic.outputVtk(vtk_root + "icostri" + std::to_string(tree_depth) + ".vtk");
      }
      {
        MeshSeed<CubedSphereSeed> quadseed;
        quadseed.setMaxAllocations(nmaxverts, nmaxedges, nmaxfaces, tree_depth);
        SpherePoisson<CubedSphereSeed> cs(nmaxverts, nmaxedges, nmaxfaces);
        cs.treeInit(tree_depth, quadseed);
        cs.updateDevice();
        cs.init();
        std::cout << "cubedsphere: ";
        cs.solve();
        cs.updateHost();
        if (!vtk_root.empty()) cs.outputVtk(vtk_root + "cubedsphere" + std::to_string(tree_depth) + ".vtk");
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "poisson_driver: " << e.what() << std::endl;
    return 1;
  }
  std::cout << "tests pass." << std::endl;
  return 0;
}
