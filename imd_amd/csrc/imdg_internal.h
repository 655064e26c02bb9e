// imdg_internal.h -- device-side data layout of the MI355X neighbour-list
// pair/EAM path.  Shared by the kernels (imdg_kernels.hip) and the host side
// of the C ABI (imdg_api.hip).
//
// Layout in HBM (one context, all replicas concatenated):
//  * atoms live in "slots"; every replica owns a run of slots padded to a
//    multiple of IMDG_BLOCK so a thread block never spans two replicas;
//    padding slots carry type -1 and take part in nothing;
//  * at every list rebuild the atoms of a replica are sorted by cell (z
//    fastest, then y, then x; inside a cell by input index), so a block of
//    IMDG_BLOCK consecutive slots is a short run of cells along z;
//  * pos[2] double4 (x, y, z, type), mom[2] double4 (px, py, pz, mass),
//    aid[2] input index: the FP64 state; the second copy is the target of
//    the rebuild permutation only;
//  * rec_a / rec_b: the gather records of the force loops, one 16 B
//    (mixed) or 32 B (FP64) record per slot: u = x - centre(cell) in the
//    force precision plus the type (rec_a, read by the density loop) or
//    2F'(rho) (rec_b, read by the force loop).  Relative to the cell centre
//    the coordinates are small, so FP32 keeps ~1e-7 A; the separation of a
//    listed pair is (u_j - u_i) + offset(cell offset), offset being
//    o . (box / grid) for o in {-1,0,1}^3 (periodic images included);
//  * tiles: for every block, the slot runs of all cells adjacent to the
//    block's cells (runs of consecutive cells along z are contiguous slot
//    ranges); the force loops copy the block's tile into LDS with coalesced
//    loads and gather neighbours from LDS;
//  * nbr: uint32 full neighbour list, entry = index | code << 27 where code
//    is the cell offset (ox+1)*9+(oy+1)*3+(oz+1) and index the neighbour's
//    position in the block's LDS tile (global slot if the tile did not fit
//    into IMDG_TILE_CAP records); padded per atom to a multiple of 4 with
//    entries pointing at the atom itself through the far-away code
//    IMDG_CODE_PAD; stored four per lane: entries 4c..4c+3 of slot i form
//    the uint4 at ((i / 64) * (kmax / 4) + c) * 64 + i % 64;
//  * tables as per-segment coefficients (p0, dv, d2v) of IMD's 3-point
//    interpolation, precomputed in FP64 and rounded to the force precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IMDG_BLOCK 128
#define IMDG_WAVE 64
#define IMDG_NWAVE (IMDG_BLOCK / IMDG_WAVE)
#define IMDG_MAXTYPES 8
#define IMDG_MAXCOLS (IMDG_MAXTYPES * IMDG_MAXTYPES)
#define IMDG_RING 4096
#define IMDG_IDX_BITS 27
#define IMDG_IDX_MASK ((1u << IMDG_IDX_BITS) - 1u)
#define IMDG_CODE_SELF 13     // cell offset (0, 0, 0)
#define IMDG_CODE_PAD 27      // list padding: offset far beyond any cutoff
#define IMDG_NCODES 28
#define IMDG_PAD_SHIFT 1.0e6
#define IMDG_GROUP 4          // list entries per lane and load
#define IMDG_TILE_CAP 2048    // records of one block's LDS tile
#define IMDG_MAXRUN 40        // slot runs of one tile

// FP64 block partials written by the force/move kernels
#define PB_EPAIR 0
#define PB_VIRIAL 1
#define PB_KE 2      // NVE: sum (k1+k2)/(4m)
#define PB_E1 3      // NVT: sum k1/m
#define PB_E2 4      // NVT: sum k2/m
#define PB_MAXD 5    // max displacement^2 (max, not sum)
#define PB_N 6

// done counters (last-block-done pattern) per kernel kind
#define DC_PERMUTE 0
#define DC_LIST 1
#define DC_FINAL 2
#define DC_N 4

struct RepDev {
  double box[3][3];       // rows: box_x, box_y, box_z
  double tbox[3][3];      // dual vectors (make_box)
  double cvec[3][3];      // cell edge vectors box_d / gdim_d
  double offs[IMDG_NCODES][3];  // o . cvec for cell offset code o;
                                // code IMDG_CODE_PAD: list padding
  double cellsz;          // list radius squared, skin included
  double skin2;           // (nbl_margin / 2)^2
  double temperature;     // NVT target kT
  double isq_tau_eta;     // 1 / tau_eta^2
  int gdim[3];
  int pbc[3];
  int atom_off;           // first slot
  int natoms;             // real atoms (slots atom_off .. atom_off+natoms-1)
  int nslots;             // padded slot count
  int cell_off;           // first cell
  int ncells;
  int blk_off;            // first block
  int nblk;
  int nactive;            // 3 * natoms
};

struct RepState {
  double eta;
  double epot, virial;    // from the last force evaluation
  int need_rebuild;
  int rebuilt;            // list rebuilt in the current step
  long nbl_count;
};

struct ThermoRec {
  double epot, ekin, virial, eta, max_disp2;
  long step;
  int rebuilt, natoms;
};

struct Ctl {
  int pos_cur;            // current position/momentum/index buffers
  int rec_cur;            // (flipped together by the rebuild permutation)
  int any_rebuild;        // some replica needs a list rebuild
  int frozen;             // list overflow: every kernel is a no-op until
                          // the host grows the list and re-runs the build
  int overflow;           // set by the list build, turned into frozen by
                          // its last block (frozen never changes mid-kernel)
  int kmax_seen;
  int is_short;
  int tiles_fallback;     // blocks whose tile exceeded IMDG_TILE_CAP
  unsigned int done[DC_N];
  long steps_done;
};

// One block's tile: slot runs copied to consecutive LDS records.
struct Tile {
  int nrun;
  int natoms;             // records in the tile
  int ok;                 // 0: did not fit, list holds global slots
  int pad;
  int4 run[IMDG_MAXRUN];  // (first slot, length, first LDS record, 0)
};

// Table of one interpolated function set on the device.
template <typename R>
struct Seg {
  R p0, dv, d2v, pad;
};

template <typename R>
struct TabDev {
  int ncols;
  int off[IMDG_MAXCOLS];
  int len[IMDG_MAXCOLS];  // usable segments (rows of the IMD column)
  R begin[IMDG_MAXCOLS], end[IMDG_MAXCOLS], invstep[IMDG_MAXCOLS];
  R twoistep[IMDG_MAXCOLS];
};

template <typename R> struct V4;
template <> struct V4<float> { typedef float4 T; };
template <> struct V4<double> { typedef double4 T; };

// Everything the per-step kernels need, passed by value.
template <typename R>
struct StepArgs {
  typedef typename V4<R>::T R4;
  const RepDev* rep;
  RepState* rst;
  Ctl* ctl;
  const int* blk_rep;     // block -> replica
  double4* pos[2];
  double4* mom[2];
  double4* xnbl;
  const int* acell;       // cell of every slot (current order)
  const signed char* atype;  // type of every slot (current order)
  R4* rec_a;              // (u, type)
  R4* rec_b;              // (u, 2F')
  const Tile* tiles;
  const uint32_t* nbr;
  const int* nnb;
  const Seg<R>* seg;      // segment coefficients of all tables
  R* rho;
  R* epot;
  R4* frc;                // (fx, fy, fz, 0)
  double* partA;          // [nblk] embedding energy partials
  double* partB;          // [nblk * PB_N]
  ThermoRec* ring;        // [IMDG_RING * nrep]
  int nrep;
  int nblk;
  int kmax;
  int ntypes;
  int ensemble;
  double dt;
  int has_pair, has_eam;
  int xcd_remap;          // map blocks to XCD-contiguous atom runs
};

struct BuildArgs {
  const RepDev* rep;
  RepState* rst;
  Ctl* ctl;
  const int* blk_rep;
  const int* cell_rep;
  double4* pos[2];
  double4* mom[2];
  int* aid[2];
  double4* xnbl;
  void* rec_a;            // float4 or double4 per slot (see mixed)
  int mixed;
  int* cellid;            // cell of slot (binning order)
  int* cslot;             // slot inside the cell from the atomic counter
  int* acell;             // cell of slot after the permutation
  signed char* atype;     // type of slot after the permutation
  int* perm;              // new slot -> old slot
  int* cell_cnt;          // atomic counters (zero outside a rebuild)
  int* cell_beg;
  int* cell_num;
  Tile* tiles;
  uint32_t* nbr;
  int* nnb;
  int nrep;
  int nslots;
  int nblk;
  int ncells;
  int kmax;
};
