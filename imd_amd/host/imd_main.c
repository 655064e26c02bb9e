/*
 * imd_main.c -- "imd_gpu [-r n] -p paramfile": IMD's program sequence for
 * the neighbour-list force path (src/imd.c:32-467): command line,
 * parameters, potentials, configuration (read, generated, or the
 * checkpoint of a restart), Maxwell velocities if needed and not
 * restarting, the .eng header unless restarting, then the main loop of
 * each simulation phase until the parameter file has no further phase or a
 * phase ends the simulation.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "imd_host.h"

int main(int argc, char **argv) {
  read_command_line(argc, argv);
  init_parameters();
  int finished = read_parameters(paramfilename, simulation);
  setup_potentials();
  if (infilename[0])
    read_atoms(infilename);
  else
    generate_fcc();
  if (!imdrestart && do_maxwell) maxwell(temperature);
  do_maxwell = 0;
  imd_gpu_init();
  if (!imdrestart && eng_int > 0) write_eng_file_header();
  if (steps_min <= steps_max) {
    printf("Starting simulation %d\n", simulation);
    if (main_loop()) finished = 1;
  }
  while (!finished) {
    simulation++;
    finished = read_parameters(paramfilename, simulation);
    if (steps_min <= steps_max) {
      imd_gpu_reconfigure();
      printf("Starting simulation %d\n", simulation);
      if (main_loop()) finished = 1;
    }
  }
  imd_gpu_finish();
  printf("Done simulation.\n");
  return 0;
}
