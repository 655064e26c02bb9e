// imdg_internal.h -- device-side data layout of the MI355X neighbour-list
// pair/EAM path.  Shared by the kernels (imdg_build.hip, imdg_step.hip) and
// the host side of the C ABI (imdg_api.hip).
//
// Layout in HBM (one context, all replicas concatenated):
//  * atoms live in "slots"; every replica owns a run of slots padded to a
//    multiple of IMDG_BLOCK; padding slots take part in nothing;
//  * at every list rebuild the atoms of a replica are sorted by cell (z
//    fastest, then y, then x; inside a cell by input index), so a chunk of
//    64 consecutive slots (one wavefront) is a short run of cells along z;
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
//  * nbr: uint32 full neighbour list, entry = slot | code << 27 where code
//    is the cell offset (ox+1)*9+(oy+1)*3+(oz+1); padded per atom to a
//    multiple of 8 with entries pointing at the atom itself through the
//    far-away code IMDG_CODE_PAD; stored four per lane: entries 4c..4c+3 of
//    slot i form the uint4 at ((i / 64) * (kmax / 4) + c) * 64 + i % 64;
//  * nbr in the 16-bit format (FP32 contexts whose loops read one-grid
//    segment tables): per 64-slot chunk, for each cell offset code, a base
//    slot (the lowest first slot of that neighbour cell over the chunk's
//    atoms); an entry is code << 11 | (slot - base[code]), 2 bytes, stored
//    in the first half of the chunk's list region: entries 4g..4g+3 of slot
//    i at halfword ((i / 64) * kmax + g * 64 + i % 64) * 4 + (entry % 4).
//    A chunk whose neighbour cells span more than 2048 slots for some code
//    (a chunk across a plane of the grid, or very dense cells) keeps the
//    32-bit format; cmeta[chunk * IMDG_META + code] holds the bases and
//    cmeta[chunk * IMDG_META + IMDG_META_WIDE] marks such chunks;
//  * tiles: for every block of IMDG_BLOCK slots, the slot runs of all
//    cells adjacent to the block's cells, staged into LDS by the list build;
//  * tables as knots (p0_k, dv_k = t[k+1] - t[k]), k = 0..len, of IMD's
//    3-point interpolation, differences taken in FP64 and then rounded to
//    the force precision; the per-step kernels copy the knots of the
//    tables their neighbour loop reads into LDS when they fit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IMDG_BLOCK 128       // replica padding, rebuild blocks
#define IMDG_WAVE 64
#define IMDG_NWAVE (IMDG_BLOCK / IMDG_WAVE)
#define IMDG_CHUNK 64        // slots per wavefront step of the step kernels
#define IMDG_STEP_THREADS 1024   // largest block of the step kernels
#define IMDG_STEP_WAVES (IMDG_STEP_THREADS / IMDG_WAVE)
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
#define IMDG_LIST_PAD 16      // lists padded to whole bodies of 4 lanes
#define IMDG_K16_BITS 11      // 16-bit entries: offset bits below the code
#define IMDG_K16_MASK ((1u << IMDG_K16_BITS) - 1u)
#define IMDG_K16_SPAN (1 << IMDG_K16_BITS)
#define IMDG_META 32          // ints of chunk metadata: bases, wide flag
#define IMDG_META_WIDE 28
// Lists whose words read per step may exceed this (slots x the longest
// list built so far, or the capacity until the host has read it, x 2 or 4
// bytes) are loaded non-temporally by the FP32 neighbour loops: half of
// MI355X's 256 MB last-level cache, which keeps a smaller list (and the
// step's other data) from one step to the next.
#define IMDG_LIST_NT_BYTES (128ull << 20)
#define IMDG_TILE_CAP 2048    // largest tile cap (records of one block's
                              // LDS tile, IMDG_TILE_LIMIT)
#define IMDG_TILE_DEFAULT 1536  // default cap: on the 2M-atom EAM crystal
                                // (tiles of ~1,400-1,900 records) the list
                                // build is fastest here, with 4 resident
                                // blocks per CU instead of 3 and about 40%
                                // of the blocks reading from HBM/L2
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
#define DC_SCAN 3
#define DC_FINBLK 4   // finalising blocks of a step done (several replicas)
#define DC_N 5

#define IMDG_SCAN_TILE 1024  // cells per block of the cell-count scan

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
  int stile_off;          // first tile of the cell-count scan
  int nstile;
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
  int tiles_fallback;     // blocks whose tile exceeded the tile cap
  int exact_pairs;        // list candidates decided by the FP64 expression
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
struct TabDev {
  int ncols;
  int off[IMDG_MAXCOLS];  // first knot of the column in the knot array
  int len[IMDG_MAXCOLS];  // usable segments (rows of the IMD column)
  R begin[IMDG_MAXCOLS], end[IMDG_MAXCOLS], invstep[IMDG_MAXCOLS];
  R twoistep[IMDG_MAXCOLS];
  // the table argument t = (x - begin) istep as t0 + x istep with the
  // constants split into R parts, istep = invstep + ilo and -begin istep =
  // t0 + t0lo (ilo = t0lo = 0 in FP64): the FP32 loops take the segment
  // from fma(x, invstep, t0) and add fma(x, ilo, t0lo) to its fraction, so
  // the rounding of the two constants shifts no fraction (a shift common
  // to every pair of a shell does not average out in the virial)
  R ilo[IMDG_MAXCOLS], t0[IMDG_MAXCOLS], t0lo[IMDG_MAXCOLS];
};

// The embedding function (and the EEAM energy term) in FP64 for the FP32
// step kernels: F and 2F' are evaluated once per atom, and an FP32
// derivative would scale the EAM part of every pair's force by the
// rounding of 2 istep, a common factor that the virial's cancellation of
// pair and EAM parts magnifies.  Knots (or cubic segments) as for the
// force-precision tables, in FP64.
struct Tab64 {
  int off[IMDG_MAXTYPES];
  int len[IMDG_MAXTYPES];
  double begin[IMDG_MAXTYPES], end[IMDG_MAXTYPES], invstep[IMDG_MAXTYPES];
  double twoistep[IMDG_MAXTYPES];
};

// Column parameters as the NT > 1 neighbour loops read them from LDS.
template <typename R>
struct ColPar {
  R begin, end, invstep, twoistep;
  int off, len;
};

template <typename R> struct V4;
template <> struct V4<float> { typedef float4 T; };
template <> struct V4<double> { typedef double4 T; };
template <typename R> struct V2;
template <> struct V2<float> { typedef float2 T; };
template <> struct V2<double> { typedef double2 T; };

// Everything the per-step kernels need, passed by value.
//
// The per-step kernels are persistent: IMDG_STEP_THREADS-thread blocks, a
// grid sized to the resident capacity, every wavefront walking its share of
// the 64-slot chunks.  Reductions go to one partial per wavefront
// (part_chunk = 0, a single replica) or one per chunk (part_chunk = 1,
// several replicas); the last block to finish sums them in a fixed order.
template <typename R>
struct StepArgs {
  typedef typename V4<R>::T R4;
  typedef typename V2<R>::T R2;
  const RepDev* rep;
  RepState* rst;
  Ctl* ctl;
  const int* blk_rep;     // IMDG_BLOCK-slot block -> replica
  double4* pos[2];
  double4* mom[2];
  double4* xnbl;
  const int* acell;       // cell of every slot (current order)
  const signed char* atype;  // type of every slot (current order)
  R4* rec_a;              // (u, type)
  R4* rec_b;              // (u, 2F')
  const uint32_t* nbr;
  const int* nnb;
  const R2* knots;        // knots of all tables
  R* rho;
  R* epot;
  R4* frc;                // (fx, fy, fz, 0)
  double* partA;          // embedding energy partials
  double* partB;          // [unit * PB_N]
  ThermoRec* ring;        // [IMDG_RING * nrep]
  int nrep;
  int nchunk;             // 64-slot chunks of all replicas
  int part_chunk;         // partials per chunk (1) or per wavefront (0)
  int unitsA;             // wavefront partials written by the density loop
  int unitsB;             // ... by the finalising force / move launch
  int store;              // write the per-atom outputs rho, epot, frc (the
                          // fused steps of a run skip them but on the last)
  int final_inline;       // the last block of that launch ends the step
                          // (0: a separate one-block launch does)
  int kmax;
  int ntypes;
  int ensemble;
  double dt;
  int has_pair, has_eam;
  int xcd_remap;          // XCD-local chunk ranges
  int lds_lo, lds_n;      // knots staged in LDS by this launch (n = 0: none)
  int joint;              // first knot of the interleaved single-species
                          // (pair, rho) table, -1 if there is none
  int dpoly;              // first knot of the FP32 rho segment polynomials
                          // (one species), -1 if there are none
  int jointn, dpolyn;     // first knot of the FP32 several-species
                          // segment records (one grid), -1 if none
  int frho, drho;         // first knot of the per-type rho segment tables
                          // the several-species loops stage in LDS when
                          // rho column it*nt + jt depends on jt only: (A,
                          // D) per segment for the force loop, (p0, A, B)
                          // as three packed floats for the density loop;
                          // -1 if none
  int pairn;              // with frho: the pair halves of the jointn
                          // records, 16 B apart (p len + k)
  int cubic;              // tables hold cubic segments, two knots each
                          // (4-point or spline interpolation)
  int eeam;               // EEAM: second density P = sum rho^2, energy E(P)
  R* eeP;                 // EEAM P per slot (written with store)
  R* eedm;                // EEAM 2E'(P) per slot (read by the force loop)
  const int* cmeta;       // chunk metadata of the 16-bit list format
  int list16;             // the list may hold 16-bit chunks
  int list_nt;            // load the list words non-temporally (a list
                          // read larger than IMDG_LIST_NT_BYTES)
  int type_bit;           // two species, FP32 shared-grid records: the
                          // neighbour's type is the lowest mantissa bit of
                          // its 2F' (rec_b.w), not a gather of atype
  const double2* knots64; // FP64 knots of the embedding (tab64[0]) and the
  const Tab64* tab64;     // EEAM energy term (tab64[1]); FP32 kernels only
  int lpa_den, lpa_frc;   // lanes per slot of the density / force loops
  int thr_den, thr_frc;   // threads per block of the density / force loops
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
  int* cell_loc;          // first slot of a cell relative to its scan tile
  int* tile_sum;          // atoms per scan tile
  int* tile_off;          // first slot of a scan tile
  const int* tile_rep;    // scan tile -> replica
  int* cell_beg;
  int* cell_num;
  Tile* tiles;
  uint32_t* nbr;
  int* nnb;
  int* blkmax;            // longest list of every block (k_list)
  int tile_cap;           // records a tile may hold (<= IMDG_TILE_CAP)
  int* cmeta;             // chunk metadata of the 16-bit list format
  int list16;             // build the 16-bit format where a chunk allows
  int nrep;
  int nslots;
  int nblk;
  int ncells;
  int nstile;
  int kmax;
};
