// Host-only driver of the library's problem compiler (fdfault_amd/csrc/
// setup.cpp) for sanitizer runs: compiles each .in file given on the command
// line with the grid built on the host and prints its time step and grid
// checksums.  Test infrastructure only.
#include <cstdio>

#include "../fdfault_amd/csrc/setup.hpp"

int main(int argc, char** argv) {
  int bad = 0;
  for (int a = 1; a < argc; a++) {
    try {
      fdf::SetupH S;
      fdf::compile_problem(argv[a], S, true, true);
      double sx = 0., sm = 0., sj = 0.;
      for (double v : S.x) sx += v;
      for (double v : S.metric) sm += v;
      for (double v : S.jac) sj += v;
      printf("%s: dt %.17g nt %d sum x %.6g metric %.6g jac %.6g\n", argv[a], S.dt, S.nt, sx, sm, sj);
    } catch (const fdf::Error& e) {
      printf("%s: error %s\n", argv[a], e.msg.c_str());
      bad = 1;
    }
  }
  return bad;
}
