/*
 * imd_oracle.h -- CPU restatement of IMD's neighbour-list pair/EAM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library (libimdg.so) never links it.
 *
 * It restates, in plain FP64 C with the reference's operation order, the
 * algorithm of these reference files (paths relative to the reference root):
 *   src/imd_potential.c:161-462   read_pot_table / formats 1 and 2
 *   src/imd_potential.c:1256-1272 init_threepoint (two guard rows)
 *   src/potaccess.h:323-354       PAIR_INT2
 *   src/potaccess.h:465-495       VAL_FUNC2
 *   src/potaccess.h:591-621       DERIV_FUNC2
 *   src/imd_geom_3d.c:52-104      make_box (tbox, heights)
 *   src/imd_geom_3d.c:113-212     init_cells (skin added once, grid sizing)
 *   src/imd_geom_3d.c:823-949     make_cell_lists (14-cell half stencil)
 *   src/imd_geom_3d.c:1054-1074   cell_coord
 *   src/imd_main_3d.c:1972-2065   do_boundaries
 *   src/imd_comm_force_3d.c:222-397, 569-716, 726-782, 897-937, 1031-1100
 *                                 buffer cells: copy_cell, add_forces,
 *                                 add_rho, copy_dF
 *   src/imd_forces_nbl.c:136-273  make_nblist
 *   src/imd_forces_nbl.c:281-1323 calc_forces (pair + EAM2 branches)
 *   src/imd_forces_nbl.c:2007-2037 check_nblist
 *   src/imd_integrate.c:32-497    move_atoms_nve
 *   src/imd_integrate.c:891-1147  move_atoms_nvt
 *
 * Parity pins (tests/test_oracle_*.py): quadratic-exactness of the 3-point
 * interpolation and its guard rows, closed-form tables, brute-force O(N^2)
 * minimum-image pair sets, analytic forces/energies of polynomial EAM tables,
 * finite-difference gradients and virial, Newton-3, perfect-lattice closed
 * forms, closed-form integrator updates.
 */
#ifndef IMD_ORACLE_H
#define IMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK          0
#define ORC_ERR_IO      1
#define ORC_ERR_FORMAT  2
#define ORC_ERR_ARG     3
#define ORC_ERR_NOMEM   4
#define ORC_ERR_GEOM    5
#define ORC_ERR_STATE   6

/* pot_table_t restated (src/types.h:416-428): values at table[k*ncols+col],
   (maxsteps+2) rows so the two guard rows always fit. */
typedef struct {
  int     ncols;
  int     maxsteps;
  double *begin, *end, *step, *invstep;
  int    *len;
  double *table;
  int     interp;  /* ORC_INTERP_* */
  double *table2;  /* spline second derivatives (pot_table_t.table2) */
} orc_table;

/* IMD's table interpolation, a build option there (src/potaccess.h:24-36):
   3-point (default, PAIR_INT2), FOURPOINT (PAIR_INT3), SPLINE
   (PAIR_INT_SP). */
#define ORC_INTERP_3POINT 0
#define ORC_INTERP_4POINT 1
#define ORC_INTERP_SPLINE 2

/* Table ingest.  default_format is used for header-less files
   (IMD: 2 in EAM2 builds, 1 otherwise; src/config.h:58-63).  radial != 0
   shifts each column so its last value is zero and raises *cellsz to the
   column end (src/imd_potential.c:359-367, 406, 439-450). */
int  orc_table_read(orc_table *pt, const char *path, int ncols, int radial,
                    int default_format, double *cellsz);
void orc_table_free(orc_table *pt);
int  orc_table_nrows(const orc_table *pt);
/* Switch a read table to another interpolation: the guard rows of
   init_fourpoint (src/imd_potential.c:1171-1187) or the second derivatives
   and guard row of init_spline (src/imd_potential.c:1197-1246; radial:
   first derivative zero at the end, otherwise natural), or back to
   init_threepoint. */
int  orc_table_set_interp(orc_table *pt, int interp, int radial);

/* Table access (src/potaccess.h) in the table's interpolation: PAIR_INT,
   VAL_FUNC, DERIV_FUNC.  inc is the column stride (ncols). */
void   orc_pair_int2(const orc_table *pt, int col, int inc, double r2,
                     double *pot, double *grad, int *is_short);
double orc_val_func2(const orc_table *pt, int col, int inc, double r2,
                     int *is_short);
double orc_deriv_func2(const orc_table *pt, int col, int inc, double r2,
                       int *is_short);

/* System. */
typedef struct orc_sys orc_sys;

#define ORC_NVE 0
#define ORC_NVT 1

/* box9 = box_x, box_y, box_z (rows).  cellsz_tab = max end over the radial
   tables (IMD's cellsz before init_cells adds the skin). */
orc_sys *orc_new(int ntypes, const double *box9, const int *pbc3,
                 double nbl_margin, double cellsz_tab, int ensemble,
                 double timestep);
void orc_free(orc_sys *s);
/* Tables are borrowed; the caller keeps them alive.  rho/embed NULL for a
   pair-only build (imd_nbl_nve). pair NULL for no pair term. */
int  orc_set_tables(orc_sys *s, const orc_table *pair, const orc_table *rho,
                    const orc_table *embed);
/* NVT parameters: temperature (kT), eta, isq_tau_eta = 1/tau_eta^2. */
void orc_set_nvt(orc_sys *s, double temperature, double eta,
                 double isq_tau_eta);
/* Atoms in caller order; positions are mapped back into the box like
   read_atoms (src/imd_io_3d.c:326).  mass is per atom. */
int  orc_set_atoms(orc_sys *s, int n, const long *id, const int *type,
                   const double *mass, const double *x, const double *p);

int  orc_calc_forces(orc_sys *s, int steps);
/* do_press_calc: per-atom stress tensor in calc_forces and move_atoms
   (STRESS_TENS, src/imd_forces_nbl.c:346-352, 558-581, 1283-1318;
   src/imd_integrate.c:411-431, 1085-1094); orc_get_press gives 6 per atom
   (xx yy zz yz zx xy, PRESSTENS), caller order */
void orc_set_press(orc_sys *s, int on);
/* Threads for a timing run on several cores (default 1: IMD's serial
   order).  With more than one, pair/EAM force loops (without stress or
   EEAM), the integrator and the skin test run split over threads and sum
   in another order, so results differ from the serial run in the last
   bits; the serial run is the checker. */
void orc_set_threads(orc_sys *s, int n);
/* EEAM (src/imd_forces_nbl.c:591-608, 1090-1095, 1180-1207): a second
   density P_i = sum_j rho_ij^2 with the energy term E(P) of the emod
   table (ntypes columns, not radial; src/imd_potential.c:82-85); set after
   orc_set_tables, NULL turns it off.  orc_get_eeam gives P and 2E'(P). */
int  orc_set_eeam(orc_sys *s, const orc_table *emod);
void orc_get_eeam(const orc_sys *s, double *P, double *dm);
void orc_get_press(const orc_sys *s, double *out);
void orc_move_atoms(orc_sys *s);
void orc_check_nblist(orc_sys *s);
void orc_invalidate_nblist(orc_sys *s);
/* nsteps of IMD's main loop (calc_forces, move_atoms, check_nblist).
   thermo (may be NULL) receives per step:
   Epot, Ekin, virial, eta, rebuilt(0/1). */
int  orc_run(orc_sys *s, int nsteps, double *thermo);

/* Per-atom state in caller order.  Any pointer may be NULL. */
void orc_get_atoms(const orc_sys *s, double *x, double *p, double *f,
                   double *epot, double *rho, double *df);
/* 0 tot_pot_energy, 1 virial, 2 tot_kin_energy, 3 eta, 4 cellsz (list),
   5 volume, 6 is_short flag, 7 nbl_count (list builds), 8 natoms,
   9 sum of |virial terms| (scale for reduced-precision tolerances) */
double orc_get_scalar(const orc_sys *s, int which);
void orc_get_grid(const orc_sys *s, int *gdim3);
/* Cell coordinates (global, 3 ints per atom) from the last rebuild. */
void orc_get_cells(const orc_sys *s, int *cell3);
/* Canonical half-pair set of the current list: 5 longs per pair
   (id_lo, id_hi, sx, sy, sz) with image shift of the id_hi atom relative
   to id_lo.  Returns the number of pairs (writes at most cap). */
long orc_get_pairs(const orc_sys *s, long *out, long cap);

#ifdef __cplusplus
}
#endif
#endif
