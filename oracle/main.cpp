// Command-line front end of the CPU oracle: `fdfault_oracle <file.in>` runs a
// whole problem like the reference executable (src/main.cpp:8-72) and writes
// the same output and front files.  Test infrastructure only.
#include <cstdio>
#include "oracle.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: fdfault_oracle <input.in>\n");
    return 1;
  }
  printf("Reading from input file %s\n", argv[1]);
  int rc = orc_run(argv[1], 0);
  if (rc != 0) return 1;
  printf("Problem finished\n");
  return 0;
}
