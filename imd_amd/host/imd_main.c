/*
 * imd_main.c -- "imd_gpu [-r n] -p paramfile": IMD's program sequence for
 * the neighbour-list force path (src/imd.c:32-467): command line,
 * parameters, potentials, configuration (read, generated, or the
 * checkpoint of a restart), Maxwell velocities if needed and not
 * restarting, the .eng header unless restarting, then the main loop of
 * each simulation phase until the parameter file has no further phase or a
 * phase ends the simulation, and IMD's closing summary (src/imd.c:310-364):
 * the average list-update interval and the seconds per step and atom, here
 * wall seconds of the main loop (the GPU run has no per-CPU time).
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "imd_host.h"

static double wall(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv) {
  double t0 = wall();
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
  double total = wall() - t0;
  printf("Done simulation.\n\n");
  printf("Neighbor list update every %d steps on average\n\n",
         steps_max / (nbl_count > 1 ? nbl_count : 1));
  printf("Did %d steps with %ld atoms.\n", steps_max + 1, natoms);
  printf("Used %f seconds wall time,\n", total);
  printf("%f seconds excluding setup time,\n", main_loop_seconds);
  double per = natoms > 0 ? main_loop_seconds / (steps_max + 1) / natoms : 0.0;
  printf("%e seconds per step and atom\n", per);
  printf("(inverse is %e).\n\n", per > 0 ? 1.0 / per : 0.0);
  return 0;
}
