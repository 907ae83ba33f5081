// The command-line program of the reference (source/main.cpp): a Simulation built from the
// flags, run to its end. One process, one GPU (-device selects it); errors end the program
// with status 1 and the message on stderr, where the reference aborts.
#include <cstdio>
#include <exception>

#include "simulation.h"

int main(int argc, char** argv)
{
  printf("=======================================================================\n");
  printf("Cubism UP 3D (velocity-pressure 3D incompressible Navier-Stokes solver)\n");
  printf("periodic block grid on one GPU\n");
  printf("=======================================================================\n");
  fflush(stdout);
  try {
    cubismup3d::Simulation sim(argc - 1, argv + 1);
    sim.run();
  } catch (const std::exception& e) {
    fprintf(stderr, "cup_simulation: %s\n", e.what());
    return 1;
  }
  return 0;
}
