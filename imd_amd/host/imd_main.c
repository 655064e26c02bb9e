/*
 * imd_main.c -- "imd_gpu -p paramfile": IMD's program sequence for the
 * neighbour-list force path (src/imd.c:32-467): parameters, potentials,
 * configuration (read or generated), Maxwell velocities if needed, the main
 * loop, the final configuration.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "imd_host.h"

int main(int argc, char **argv) {
  const char *pfile = NULL;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "-p") && i + 1 < argc)
      pfile = argv[++i];
    else if (!strncmp(argv[i], "-p", 2) && argv[i][2])
      pfile = argv[i] + 2;
  }
  if (!pfile) {
    fprintf(stderr, "usage: %s -p paramfile\n", argv[0]);
    return 2;
  }
  init_parameters();
  read_parameters(pfile, 1);
  check_parameters_complete();
  setup_potentials();
  if (infilename[0])
    read_atoms(infilename);
  else
    generate_fcc();
  if (do_maxwell) maxwell(temperature);
  imd_gpu_init();
  write_eng_file_header();
  main_loop();
  fetch_atoms();
  char fn[IMD_STRLEN + 16];
  snprintf(fn, sizeof fn, "%s.end.chkpt", outfilename);
  write_config(fn);
  imd_gpu_finish();
  return 0;
}
