// main.cpp -- the NS3D command line (navierstokes3D.cxx:480-545): read the
// .ini file, build NavierStokesBase on the GPU, run the requested driver.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/ns3d_host.h"
#include "NSBase.h"
#include "drivers.h"
#include "vtk_io.h"

using namespace ns3d_host;

static void PrintUsage() {
  printf("Usage: NS3D <input-file>   (or 'NS3D -h' for this help)\n");
  printf("\nInput file format: one 'key = value' per line.\n");
  printf("Keys are case-insensitive, blank lines and lines starting\n");
  printf("with '#' are ignored, and an inline '# ...' comment is\n");
  printf("allowed after a value. Unknown keys are reported as a warning.\n");
  printf("\nAccepted keys:\n");
  printf("  input          image file (binary 3D image, .vti or .vtk)   [required]\n");
  printf("  pressurefield  initial pressure field file ('none' to skip)\n");
  printf("  vxfield        initial x-velocity field file ('none' to skip)\n");
  printf("  vyfield        initial y-velocity field file ('none' to skip)\n");
  printf("  vzfield        initial z-velocity field file ('none' to skip)\n");
  printf("  dx, dy, dz     grid spacing (m); dy/dz < 0 means 'same as dx'\n");
  printf("  mu             dynamic viscosity (Pa.s)\n");
  printf("  rho            density (kg/m^3)\n");
  printf("  p0             inlet pressure (Pa)\n");
  printf("  gradp          imposed pressure gradient (Pa/m)\n");
  printf("  alpha          velocity relaxation parameter\n");
  printf("  beta           pressure relaxation parameter\n");
  printf("  vsolvertol     velocity solver tolerance\n");
  printf("  psolvertol     pressure solver tolerance\n");
  printf("  pcsolvertol    pressure correction solver tolerance\n");
  printf("  simpletol      outer-loop convergence threshold (relative RMR)\n");
  printf("  deltaRMR       oscillation detection threshold\n");
  printf("  maxitsolver    max iterations of the linear solvers\n");
  printf("  maxitSIMPLE    max outer iterations\n");
  printf("  N1             inner pressure iterations (SIMPLER/SIMPLERR/IDEAL)\n");
  printf("  N2             inner final iterations (IDEAL)\n");
  printf("  save_step      save a field snapshot every N outer iterations\n");
  printf("  porevalue      voxel value of the fluid (0 or 1)\n");
  printf("  direction      pressure gradient direction: 1=x, 2=y, 3=z\n");
  printf("  border         solid border thickness added along x (voxels)\n");
  printf("  method         algorithm: SIMPLEC, SIMPLER, SIMPLERR or IDEAL\n");
}

int main(int argc, char* argv[]) {
  if (argc < 2) {
    PrintUsage();
    return 1;
  }
  const std::string ini = argv[1];
  if (ini == "-h" || ini == "--help") {
    PrintUsage();
    return 0;
  }
  ns3d_run_params p;
  std::vector<char> names(8192, 0);
  if (ns3d_read_ini(ini.c_str(), &p, names.data(), (int)names.size()) != 0) return 1;
  std::string fields[5];
  const char* q = names.data();
  for (auto& f : fields) {
    f = q;
    q += f.size() + 1;
  }
  try {
    int nx, ny, nz;
    Image im = read_image(fields[0]);
    nx = im.nx; ny = im.ny; nz = im.nz;
    std::vector<uint8_t> raw((size_t)nx * ny * nz);
    for (size_t i = 0; i < raw.size(); ++i) raw[i] = (uint8_t)im.data[i * im.ncomp];
    printf("Dims: x: %d y: %d z: %d\n", nx, ny, nz);
    NavierStokesBase base(raw, nx, ny, nz, fields[0], fields[1], fields[2], fields[3], fields[4], p.dx, p.dy, p.dz,
                          p.vsolvertol, p.psolvertol, p.pcsolvertol, p.maxitsolver, p.alpha, p.beta, p.p0,
                          p.gradp, p.mu, p.rho, p.porevalue, p.direction, p.deltaRMR, p.save_step, p.border,
                          -1, 1, true, "");
    if (p.method == NS3D_METHOD_IDEAL) {
      IDEALSOLVER solver(p.simpletol, p.maxitSIMPLE, p.N1, p.N2);
      solver.Solve(base);
    } else if (p.method == NS3D_METHOD_SIMPLER) {
      SIMPLERSOLVER solver(p.simpletol, p.maxitSIMPLE, p.N1);
      solver.Solve(base);
    } else if (p.method == NS3D_METHOD_SIMPLERR) {
      SIMPLERRSOLVER solver(p.simpletol, p.maxitSIMPLE, p.N1);
      solver.Solve(base);
    } else {
      SIMPLECSOLVER solver(p.simpletol, p.maxitSIMPLE);
      solver.Solve(base);
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
