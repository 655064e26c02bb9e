// imdg_step.hip -- CDNA4 (gfx950) per-step kernels of the neighbour-list
// pair/EAM path: host density with embedding, pair + EAM forces fused with
// the leapfrog step and the skin test, the unfused move, compensated FP64
// reductions finished by the last block of each step.
//
// Execution model: persistent kernels.  A grid of IMDG_STEP_THREADS-thread
// blocks sized to what the chip holds at once copies the knots of the
// tables its neighbour loop reads into LDS (when they fit), then every
// wavefront walks its share of the 64-slot chunks: lane = atom, neighbour
// records gathered from HBM/L2 through the grouped full list (four entries
// per lane and load, the next groups prefetched), table knots from LDS.
// There is no block barrier after the staging, so waves never wait for
// each other.
//
// Reference semantics (paths relative to the reference root):
//   density/embedding  src/imd_forces_nbl.c:586-611, 1079-1112
//   pair + EAM force   src/imd_forces_nbl.c:490-584, 1117-1322
//   tables             PAIR_INT2/VAL_FUNC2/DERIV_FUNC2 src/potaccess.h:323-354,
//                      465-495, 591-621
//   integrators        move_atoms_nve src/imd_integrate.c:122-358,
//                      move_atoms_nvt src/imd_integrate.c:949-1141
//   skin test          check_nblist src/imd_forces_nbl.c:2007-2037
#include "imdg_device.h"

namespace imdg {

// Largest block of the force loop (its launch bound).  FP32: 1024 threads
// (16 wavefronts; the loop fits the 128 VGPRs that leaves each); FP64: 512
// (its doubles need more registers).  With the tables in LDS one block per
// CU is resident, so the block size is the resident wavefront count; the
// launch may use fewer threads (imdg_set_tuning).
template <typename R, int MODE>
struct ForceThreads {
  static constexpr int value = sizeof(R) == 4 ? 1024 : 512;
};

// Neighbour-loop variants of the force kernel.
enum {
  FM_NT1 = 0,  // one species
  FM_SG = 1,   // one species, pair and rho tables on the same grid: one
               // segment lookup and one interleaved knot table
  FM_NTN = 2,  // several species
  FM_SGN = 3   // several species, every pair and rho column on one grid
               // (FP32): one segment lookup, one 32 B record per type pair
};

// ---------------------------------------------------------------------------
// schedule, offsets, partials
// ---------------------------------------------------------------------------

// The list words (read once per launch, never gathered) are loaded
// non-temporally when the list is larger than the last-level cache can
// keep between steps (list_nt; the FP32 one-grid loops have an instance
// for each policy): they then pass the L2 without pushing out the neighbour
// records that the loops gather many times (2M-atom crystal, 0.3 GB of
// list: force 258 -> 252 us, density 194 -> 189 us).  A smaller list is
// read again from the last-level cache every step, and non-temporal loads
// would send it to HBM (256,000 atoms: 9% slower).  Positions and momenta
// keep the default policy: marked non-temporal they reach the next step's
// launch from HBM (force 252 -> 265 us).
typedef unsigned int nt_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int nt_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 nt_load2(const void* p) {
  nt_u2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u2*>(p));
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 nt_load4(const void* p) {
  nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Chunks of one wavefront: c, c + stride, ... below c1.  Blocks are
// dispatched round-robin over the 8 XCDs; with xcd_remap the waves of XCD x
// walk the x-th eighth of the chunks, so the neighbour records they gather
// stay in that XCD's L2.  unit: the wavefront's partial slot.
struct Sched {
  int c, c1, stride, unit;
};

__device__ __forceinline__ Sched make_sched(int nchunk, int xcd_remap) {
  Sched s;
  int nwb = blockDim.x / IMDG_WAVE, w = threadIdx.x / IMDG_WAVE;
  s.unit = blockIdx.x * nwb + w;
  if (xcd_remap && (gridDim.x & 7) == 0) {
    int x = blockIdx.x & 7, bx = blockIdx.x >> 3;
    int per = (nchunk + 7) >> 3;
    int c0 = min(x * per, nchunk);
    s.c1 = min(c0 + per, nchunk);
    s.stride = (gridDim.x >> 3) * nwb;
    s.c = c0 + bx * nwb + w;
  } else {
    s.c1 = nchunk;
    s.stride = gridDim.x * nwb;
    s.c = s.unit;
  }
  return s;
}

// Cell-offset vectors of a replica, in the force precision, into the
// wavefront's own LDS row (one 4-vector per code: one LDS load per pair),
// and twice what rounding to the force precision dropped from them
// (olo).  Every pair with the same cell offset would otherwise see the same
// error in its separation, a small uniform strain per direction that the
// virial's cancellation of large terms magnifies; the FP32 one-grid loops
// fold 2 d.olo into r^2 (sep_r2).
template <typename R4>
__device__ __forceinline__ void wave_offsets(const RepDev& Rp, R4* off,
                                             R4* olo) {
  typedef decltype(off->x) R;
  int l = threadIdx.x & (IMDG_WAVE - 1);
  if (l < IMDG_NCODES) {
    R4 o, q;
    o.x = (R)Rp.offs[l][0];
    o.y = (R)Rp.offs[l][1];
    o.z = (R)Rp.offs[l][2];
    o.w = 0;
    q.x = (R)(2.0 * (Rp.offs[l][0] - (double)o.x));
    q.y = (R)(2.0 * (Rp.offs[l][1] - (double)o.y));
    q.z = (R)(2.0 * (Rp.offs[l][2] - (double)o.z));
    q.w = 0;
    off[l] = o;
    olo[l] = q;
  }
  wave_lds_sync();
}

// r^2 of a separation d formed with the rounded cell offset, with 2 d.olo
// (olo: twice the offset's rounding remainder) added inside the FMA chain,
// where the partial sums still resolve it: the rounding of r^2 is then
// unbiased over the pairs of a shell.
__device__ __forceinline__ float sep_r2(float dx, float dy, float dz,
                                        const float4& ol) {
  float c = dx * ol.x;
  c = fmaf(dy, ol.y, c);
  c = fmaf(dz, ol.z, c);
  return fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, c)));
}

// Wavefront reduction of the step partials (sum; max for PB_MAXD).
__device__ __forceinline__ void wave_reduce_part(double (&v)[PB_N]) {
#pragma unroll
  for (int q = 0; q < PB_N; q++)
    v[q] = (q == PB_MAXD) ? wave_max(v[q]) : wave_sum(v[q]);
}

__device__ __forceinline__ void add_part(double (&acc)[PB_N],
                                         const double (&v)[PB_N]) {
#pragma unroll
  for (int q = 0; q < PB_N; q++)
    acc[q] = (q == PB_MAXD) ? fmax(acc[q], v[q]) : acc[q] + v[q];
}

// Partials of one chunk (part_chunk) or accumulation into the wavefront's
// running sums.
__device__ __forceinline__ void put_part(double* partB, int part_chunk,
                                         int c, double (&v)[PB_N],
                                         double (&acc)[PB_N]) {
  if (part_chunk) {
    wave_reduce_part(v);
    if ((threadIdx.x & (IMDG_WAVE - 1)) == 0) {
#pragma unroll
      for (int q = 0; q < PB_N; q++) partB[(size_t)c * PB_N + q] = v[q];
    }
  } else {
    add_part(acc, v);
  }
}

// Sums of NQ per-wavefront values over the wavefronts of the block, in
// wavefront order with TwoSum (the result does not depend on timing),
// written as the block's partial out[q] by thread q.  Max instead of sum
// for q == qmax.  Called by every thread of the block, after its loops.
template <int NQ>
__device__ __forceinline__ void block_partial(double* out, const double* v,
                                              int qmax) {
  __shared__ double lp[IMDG_STEP_WAVES][NQ];
  int w = threadIdx.x / IMDG_WAVE, nw = blockDim.x / IMDG_WAVE;
  if ((threadIdx.x & (IMDG_WAVE - 1)) == 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) lp[w][q] = v[q];
  }
  __syncthreads();
  int q = threadIdx.x;
  if (q < NQ) {
    double s = lp[0][q], c = 0.0;
    for (int k = 1; k < nw; k++) {
      if (q == qmax) {
        s = fmax(s, lp[k][q]);
      } else {
        double t, e;
        two_sum(s, lp[k][q], t, e);
        s = t;
        c += e;
      }
    }
    out[q] = s + c;
  }
}

// One replica: the wavefronts' running sums, reduced over the block into
// the block's partial (partB[block]).
__device__ __forceinline__ void flush_part(double* partB, int part_chunk,
                                           double (&acc)[PB_N]) {
  if (part_chunk) return;
  wave_reduce_part(acc);
  block_partial<PB_N>(partB + (size_t)blockIdx.x * PB_N, acc, PB_MAXD);
}

__device__ __forceinline__ void flag_short(Ctl* ctl, int is_short) {
  if (__ballot(is_short) != 0 && (threadIdx.x & (IMDG_WAVE - 1)) == 0)
    atomicOr(&ctl->is_short, 1);
}

// Knots of the launch's LDS range (block-wide, before any wave loop).
template <typename R2>
__device__ __forceinline__ void stage_knots(const R2* __restrict__ g, int lo,
                                            int n, R2* lds) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) lds[k] = g[lo + k];
}

template <typename R>
__device__ __forceinline__ void stage_colpar(const TabDev<R>& t, int klo,
                                             ColPar<R>* cp) {
  for (int k = threadIdx.x; k < t.ncols; k += blockDim.x) {
    ColPar<R> p;
    p.begin = t.begin[k];
    p.end = t.end[k];
    p.invstep = t.invstep[k];
    p.twoistep = t.twoistep[k];
    p.off = t.off[k] - klo;
    p.len = t.len[k];
    cp[k] = p;
  }
}

// ---------------------------------------------------------------------------
// finalisation (last block of a step kernel)
// ---------------------------------------------------------------------------

// Per-replica end of a step from the replica's summed partials (tot[PB_N]
// is the embedding-energy sum): energies, thermostat, skin decision, thermo
// record.  Returns whether the replica needs a list rebuild.
template <typename R>
__device__ bool replica_step_end(StepArgs<R>& a, int r, const double* tot,
                                 bool forces, bool moved) {
  const RepDev& Rp = a.rep[r];
  RepState& st = a.rst[r];
  if (forces) {
    st.epot = tot[PB_EPAIR] + tot[PB_N];
    st.virial = tot[PB_VIRIAL];
  }
  if (!moved) return false;
  double ekin;
  if (a.ensemble == 0) {
    ekin = tot[PB_KE];
  } else {
    ekin = (tot[PB_E1] + tot[PB_E2]) / 4.0;
    double ttt = Rp.nactive * Rp.temperature;
    // a replica without atoms has no temperature to steer
    if (Rp.nactive > 0)
      st.eta += a.dt * (tot[PB_E2] / ttt - 1.0) * Rp.isq_tau_eta;
  }
  int need = tot[PB_MAXD] > Rp.skin2;
  st.need_rebuild = need;
  long step = a.ctl->steps_done;
  ThermoRec& tr = a.ring[(size_t)(step % IMDG_RING) * a.nrep + r];
  tr.epot = st.epot;
  tr.ekin = ekin;
  tr.virial = st.virial;
  tr.eta = st.eta;
  tr.max_disp2 = tot[PB_MAXD];
  tr.step = step;
  tr.rebuilt = st.rebuilt;
  tr.natoms = Rp.natoms;
  st.rebuilt = 0;
  return need != 0;
}

// Wavefront-wide compensated combine of per-lane (sum, compensation)
// pairs in a fixed shuffle order (TwoSum keeps the error); max for
// PB_MAXD.  Result valid in every lane.
__device__ __forceinline__ void wave_combine(double (&s)[PB_N + 1],
                                             double (&c)[PB_N + 1]) {
#pragma unroll
  for (int q = 0; q <= PB_N; q++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double so = __shfl_xor(s[q], o, IMDG_WAVE);
      double co = __shfl_xor(c[q], o, IMDG_WAVE);
      if (q == PB_MAXD) {
        s[q] = fmax(s[q], so);
      } else {
        double t, e;
        two_sum(s[q], so, t, e);
        s[q] = t;
        c[q] = c[q] + co + e;
      }
    }
  }
}

// Compensated sums of the partials of every replica (fixed order, so the
// result does not depend on which block finishes last), thermostat, skin
// decision, thermo record.  One replica: the whole block sums the block
// partials (fblk = 0, nfblk = 1).  Several replicas: one wavefront per
// replica sums that replica's chunk partials, with no block barrier per
// replica; the replicas are spread over the nfblk finalising blocks (this
// is block fblk), which OR their rebuild decisions into the control block.
template <typename R>
__device__ void finalize_step(StepArgs<R>& a, bool forces, bool moved,
                              int fblk, int nfblk) {
  __shared__ double lred[IMDG_STEP_WAVES][PB_N + 1][2];
  __shared__ int any_s;
  int nw = blockDim.x / IMDG_WAVE;
  int w = threadIdx.x / IMDG_WAVE, l = threadIdx.x % IMDG_WAVE;
  if (threadIdx.x == 0) any_s = 0;
  __syncthreads();
  if (!a.part_chunk) {
    // one replica: partials per wavefront of the finalising kernel (B) and
    // of the density kernel (A)
    int unitsB = a.unitsB;
    double s[PB_N + 1], c[PB_N + 1];
#pragma unroll
    for (int q = 0; q <= PB_N; q++) s[q] = c[q] = 0.0;
    for (int u = threadIdx.x; u < unitsB; u += blockDim.x) {
      const double* pb = a.partB + (size_t)u * PB_N;
#pragma unroll
      for (int q = 0; q < PB_N; q++) {
        if (q == PB_MAXD)
          s[q] = fmax(s[q], pb[q]);
        else
          kahan_add(s[q], c[q], pb[q]);
      }
    }
    if (forces && a.has_eam)
      for (int u = threadIdx.x; u < a.unitsA; u += blockDim.x)
        kahan_add(s[PB_N], c[PB_N], a.partA[u]);
    wave_combine(s, c);
    if (l == 0) {
#pragma unroll
      for (int q = 0; q <= PB_N; q++) {
        lred[w][q][0] = s[q];
        lred[w][q][1] = c[q];
      }
    }
    __syncthreads();
    // the wavefront results in wavefront order, one quantity per thread
    __shared__ double ltot[PB_N + 1];
    int q = threadIdx.x;
    if (q <= PB_N) {
      double ss = lred[0][q][0], cc = lred[0][q][1];
      for (int k = 1; k < nw; k++) {
        if (q == PB_MAXD) {
          ss = fmax(ss, lred[k][q][0]);
        } else {
          double t, e;
          two_sum(ss, lred[k][q][0], t, e);
          ss = t;
          cc = cc + lred[k][q][1] + e;
        }
      }
      ltot[q] = (q == PB_MAXD) ? ss : ss + cc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot[PB_N + 1];
#pragma unroll
      for (int k = 0; k <= PB_N; k++) tot[k] = ltot[k];
      if (replica_step_end(a, 0, tot, forces, moved)) any_s = 1;
    }
  } else {
    // several replicas: one wavefront per replica, chunk partials
    for (int r = fblk * nw + w; r < a.nrep; r += nfblk * nw) {
      const RepDev& Rp = a.rep[r];
      int c0 = Rp.blk_off * (IMDG_BLOCK / IMDG_CHUNK);
      int c1 = (Rp.blk_off + Rp.nblk) * (IMDG_BLOCK / IMDG_CHUNK);
      double s[PB_N + 1], c[PB_N + 1];
#pragma unroll
      for (int q = 0; q <= PB_N; q++) s[q] = c[q] = 0.0;
      for (int u = c0 + l; u < c1; u += IMDG_WAVE) {
        const double* pb = a.partB + (size_t)u * PB_N;
#pragma unroll
        for (int q = 0; q < PB_N; q++) {
          if (q == PB_MAXD)
            s[q] = fmax(s[q], pb[q]);
          else
            kahan_add(s[q], c[q], pb[q]);
        }
        if (forces && a.has_eam) kahan_add(s[PB_N], c[PB_N], a.partA[u]);
      }
      wave_combine(s, c);
      if (l == 0) {
        double tot[PB_N + 1];
#pragma unroll
        for (int q = 0; q <= PB_N; q++)
          tot[q] = (q == PB_MAXD) ? s[q] : s[q] + c[q];
        if (replica_step_end(a, r, tot, forces, moved)) atomicOr(&any_s, 1);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (moved && nfblk == 1) {
      a.ctl->any_rebuild = any_s ? 1 : 0;
      a.ctl->steps_done += 1;
    } else if (moved) {
      // every block has read steps_done (its thermo records) before it
      // counts itself done; the last one advances the step.  The flag is
      // clear at this point (the last rebuild cleared it).
      if (any_s) atomicOr(&a.ctl->any_rebuild, 1);
      unsigned t = atomicAdd(&a.ctl->done[DC_FINBLK], 1u);
      if (t == (unsigned)nfblk - 1) {
        a.ctl->done[DC_FINBLK] = 0;
        a.ctl->steps_done += 1;
      }
    }
    __threadfence();
  }
}

template <typename R>
__device__ __forceinline__ void last_block_finalize(StepArgs<R>& a,
                                                    bool forces, bool moved) {
  __shared__ int is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(&a.ctl->done[DC_FINAL], 1u);
    is_last = (t == gridDim.x - 1);
    if (is_last) a.ctl->done[DC_FINAL] = 0;
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    finalize_step(a, forces, moved, 0, 1);
  }
}

// ---------------------------------------------------------------------------
// integration
// ---------------------------------------------------------------------------

// Kinetic terms of one atom's step: NVE ke = (|p|^2 before + after) / 4m;
// NVT e1, e2 = |p|^2 / m before and after.
struct KinTerms {
  double ke, e1, e2;
};

// Leapfrog on momenta (src/imd_integrate.c:213-217, 329-358 and
// 1020-1056), FP64, separately rounded like the reference.  The NVT
// factors reib and einsd are the same for every atom of the replica.
__device__ __forceinline__ KinTerms leapfrog(int ensemble, double dt,
                                             double reib, double einsd,
                                             double fx, double fy, double fz,
                                             double4& p, double& x0,
                                             double& x1, double& x2) {
#pragma clang fp contract(off)
  KinTerms kt;
  double m = p.w;
  if (ensemble == 0) {
    double k1 = p.x * p.x + p.y * p.y + p.z * p.z;
    p.x += dt * fx;
    p.y += dt * fy;
    p.z += dt * fz;
    double k2 = p.x * p.x + p.y * p.y + p.z * p.z;
    kt.ke = (k1 + k2) / (4 * m);
    kt.e1 = kt.e2 = 0.0;
  } else {
    kt.e1 = (p.x * p.x + p.y * p.y + p.z * p.z) / m;
    p.x = (p.x * reib + dt * fx) * einsd;
    p.y = (p.y * reib + dt * fy) * einsd;
    p.z = (p.z * reib + dt * fz) * einsd;
    kt.e2 = (p.x * p.x + p.y * p.y + p.z * p.z) / m;
    kt.ke = 0.0;
  }
  double tmp = dt / m;
  x0 += tmp * p.x;
  x1 += tmp * p.y;
  x2 += tmp * p.z;
  return kt;
}

// reib = 1 - eta dt/2, einsd = 1 / (1 + eta dt/2) of a replica
// (src/imd_integrate.c:907-908).
__device__ __forceinline__ void nvt_factors(double eta, double dt,
                                            double& reib, double& einsd) {
#pragma clang fp contract(off)
  reib = 1.0 - eta * dt / 2.0;
  einsd = 1.0 / (1.0 + eta * dt / 2.0);
}

// Integrate slot i, refresh its cell-relative record and return the
// partials (kinetic terms, displacement).
template <typename R>
__device__ __forceinline__ void integrate_slot(StepArgs<R>& a,
                                               const RepDev& Rp, int r, int i,
                                               double fx, double fy, double fz,
                                               double (&v)[PB_N]) {
  typedef typename V4<R>::T R4;
  int pc = a.ctl->pos_cur;
  double4* pos = a.pos[pc];
  double4* mom = a.mom[pc];
  double4 xi = pos[i];
  double4 p = mom[i];
  double x0 = xi.x, x1 = xi.y, x2 = xi.z;
  double reib = 1.0, einsd = 1.0;
  if (a.ensemble != 0) nvt_factors(a.rst[r].eta, a.dt, reib, einsd);
  KinTerms kt =
      leapfrog(a.ensemble, a.dt, reib, einsd, fx, fy, fz, p, x0, x1, x2);
  mom[i] = p;
  double4 xn = make_double4(x0, x1, x2, xi.w);
  pos[i] = xn;
  double c[3];
  cell_centre(Rp, a.acell[i], c);
  R4 ra;
  ra.x = (R)(x0 - c[0]);
  ra.y = (R)(x1 - c[1]);
  ra.z = (R)(x2 - c[2]);
  ra.w = (R)xi.w;
  a.rec_a[i] = ra;
  v[PB_KE] = kt.ke;
  v[PB_E1] = kt.e1;
  v[PB_E2] = kt.e2;
  v[PB_MAXD] = disp2(xn, a.xnbl[i]);
}


// ---------------------------------------------------------------------------
// neighbour-list walk
// ---------------------------------------------------------------------------

// Lanes per atom of the neighbour loops.  With L > 1 a wavefront walks its
// 64-slot chunk in L passes of 64/L slots; the L lanes of a slot take
// interleaved entries of its list, so the records one gather instruction
// reads for a slot are L consecutive list entries -- mostly neighbouring
// slots of one cell, one or two cache lines instead of L.  Pass p ends with
// a sum over the slot's L lanes, and lane p of the slot keeps it: after the
// L passes lane l holds the sums of chunk slot epi_slot(l).  The per-atom
// work (embedding, integration, writes) then runs with lane = slot, so its
// loads and stores are contiguous: lane l takes the sums of slot l from
// lane keep_lane(l) (a cross-lane shuffle per sum).
__device__ __forceinline__ int epi_slot(int lane, int L) {
  return (IMDG_WAVE / L) * (lane % L) + lane / L;
}
__device__ __forceinline__ int keep_lane(int slot, int L) {
  return (slot % (IMDG_WAVE / L)) * L + slot / (IMDG_WAVE / L);
}
template <int L, typename R>
__device__ __forceinline__ R to_slot_lane(R v, int lane) {
  return L == 1 ? v : __shfl(v, keep_lane(lane, L), IMDG_WAVE);
}

// Sum over the L lanes of a slot (aligned lane groups); every lane of the
// group gets the same bits (each butterfly step adds the same two values).
template <int L, typename R>
__device__ __forceinline__ R lane_group_sum(R v) {
  if (L >= 2) v += __shfl_xor(v, 1, IMDG_WAVE);
  if (L >= 4) v += __shfl_xor(v, 2, IMDG_WAVE);
  return v;
}
template <int L, typename R>
__device__ __forceinline__ R lane_group_min(R v) {
  if (L >= 2) v = rmin(v, __shfl_xor(v, 1, IMDG_WAVE));
  if (L >= 4) v = rmin(v, __shfl_xor(v, 2, IMDG_WAVE));
  return v;
}

// Four list entries of one lane: with L lanes per slot, one body covers L
// groups of the grouped layout (4L entries of the slot); lane s of the slot
// reads words s*(4/L) .. s*(4/L) + 4/L - 1 of each of them.
template <int L, bool NT = false>
struct ListRef {
  static constexpr bool W16 = false;
  const uint32_t* p;  // the lane's first word of the body
  __device__ __forceinline__ uint4 load() const {
    constexpr int G = IMDG_WAVE * IMDG_GROUP;  // words per group
    if (!NT) {  // NT: non-temporal loads (the list_nt policy)
      if (L == 1) return *reinterpret_cast<const uint4*>(p);
      if (L == 2) {
        uint2 x = *reinterpret_cast<const uint2*>(p);
        uint2 y = *reinterpret_cast<const uint2*>(p + G);
        return make_uint4(x.x, x.y, y.x, y.y);
      }
      return make_uint4(p[0], p[G], p[2 * G], p[3 * G]);
    }
    if (L == 1) return nt_load4(p);
    if (L == 2) {
      uint2 x = nt_load2(p);
      uint2 y = nt_load2(p + G);
      return make_uint4(x.x, x.y, y.x, y.y);
    }
    return make_uint4(__builtin_nontemporal_load(p),
                      __builtin_nontemporal_load(p + G),
                      __builtin_nontemporal_load(p + 2 * G),
                      __builtin_nontemporal_load(p + 3 * G));
  }
};

// First word of lane s's share of slot i's list.
template <int L>
__device__ __forceinline__ const uint32_t* list_words(const uint32_t* nbr,
                                                      int i, int kmax, int s) {
  return reinterpret_cast<const uint32_t*>(list_of(nbr, i, kmax)) +
         s * (IMDG_GROUP / L);
}

// Marks the four list words as needed here: the memory-counter wait for
// the load that fills them is placed at this point (an empty statement).
__device__ __forceinline__ void ready(const uint4& e) {
  asm volatile("" ::"v"(e.x), "v"(e.y), "v"(e.z), "v"(e.w));
}

// Walk a slot's list, one body (four entries per lane) at a time, in pairs
// A = 2h, B = 2h+1.  A and B sit in two registers of their own; each body
// issues its record gathers, then (through pf/slot) the load of the body two
// ahead into the register it has just consumed, then computes.  No register
// changes role between iterations, so nothing waits for the prefetch
// before the next body's gathers need it.  The initial values are selects
// rather than plain loads: a phi of loads would be folded into one load
// next to its use, which undoes the prefetch.  Lists are padded with
// far-away entries to a multiple of 16, so with 1 or 2 lanes per slot the
// body count is rounded up to even inside the padding and B always runs;
// with 4 lanes a body is 16 entries and B past the last body is skipped.
template <int L, bool NT = false, typename F>
__device__ __forceinline__ void list_walk(const uint32_t* w, int nnb, F body) {
  constexpr int E = IMDG_GROUP * L;                    // entries per body
  constexpr size_t BW = (size_t)L * IMDG_WAVE * IMDG_GROUP;  // words per body
  int nbody = (nnb + E - 1) / E;
  if (L <= 2) nbody = (nbody + 1) & ~1;
  const int blast = nbody > 0 ? nbody - 1 : 0;
  uint4 A = ListRef<L, NT>{w}.load();
  uint4 B = ListRef<L, NT>{w + (size_t)min(1, blast) * BW}.load();
  A = make_uint4(nbody > 0 ? A.x : 0u, nbody > 0 ? A.y : 0u,
                 nbody > 0 ? A.z : 0u, nbody > 0 ? A.w : 0u);
  B = make_uint4(nbody > 0 ? B.x : 0u, nbody > 0 ? B.y : 0u,
                 nbody > 0 ? B.z : 0u, nbody > 0 ? B.w : 0u);
  ready(A);  // both ways into the loop head then have A's words in hand
  for (int h = 0; 2 * h < nbody; h++) {
    body(A, ListRef<L, NT>{w + (size_t)min(2 * h + 2, blast) * BW}, A);
    if (L <= 2 || 2 * h + 1 < nbody) {
      body(B, ListRef<L, NT>{w + (size_t)min(2 * h + 3, blast) * BW}, B);
      // take the wait for the next A here, where only B's prefetch is newer
      // than it (at the loop head the wait counter is merged over the loop
      // entry and would also wait for B's prefetch)
      ready(A);
    }
  }
}

// The 16-bit format (a chunk's entries code << 11 | slot - base[code],
// halfword groups of 64 lanes): the same lane pattern, each load half as
// wide.  load() returns the loaded words as they are (so the wait for them
// stays where the walk places it); unpack16 spreads them to four entries.
template <int L, bool NT = false>
struct ListRef16 {
  static constexpr bool W16 = true;
  const uint16_t* p;  // the lane's first halfword of the body
  __device__ __forceinline__ uint4 load() const {
    constexpr int G = IMDG_WAVE * IMDG_GROUP;  // halfwords per group
    if (!NT) {  // NT: non-temporal loads (the list_nt policy)
      if (L == 1) {
        uint2 x = *reinterpret_cast<const uint2*>(p);
        return make_uint4(x.x, x.y, 0u, 0u);
      }
      if (L == 2) {
        uint32_t x = *reinterpret_cast<const uint32_t*>(p);
        uint32_t y = *reinterpret_cast<const uint32_t*>(p + G);
        return make_uint4(x, y, 0u, 0u);
      }
      return make_uint4(p[0], p[G], p[2 * G], p[3 * G]);
    }
    if (L == 1) {
      uint2 x = nt_load2(p);
      return make_uint4(x.x, x.y, 0u, 0u);
    }
    if (L == 2) {
      uint32_t x =
          __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
      uint32_t y =
          __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p + G));
      return make_uint4(x, y, 0u, 0u);
    }
    return make_uint4(__builtin_nontemporal_load(p),
                      __builtin_nontemporal_load(p + G),
                      __builtin_nontemporal_load(p + 2 * G),
                      __builtin_nontemporal_load(p + 3 * G));
  }
};

template <int L>
__device__ __forceinline__ uint4 unpack16(uint4 r) {
  if (L == 4) return r;
  return make_uint4(r.x & 0xffffu, r.x >> 16, r.y & 0xffffu, r.y >> 16);
}

template <int L>
__device__ __forceinline__ const uint16_t* list_words16(const uint32_t* nbr,
                                                        int i, int kmax,
                                                        int s) {
  return reinterpret_cast<const uint16_t*>(list16_of(nbr, i, kmax)) +
         s * (IMDG_GROUP / L);
}

// list_walk on the 16-bit format.
template <int L, bool NT = false, typename F>
__device__ __forceinline__ void list_walk16(const uint16_t* w, int nnb,
                                            F body) {
  constexpr int E = IMDG_GROUP * L;
  constexpr size_t BW = (size_t)L * IMDG_WAVE * IMDG_GROUP;  // halfwords
  int nbody = (nnb + E - 1) / E;
  if (L <= 2) nbody = (nbody + 1) & ~1;
  const int blast = nbody > 0 ? nbody - 1 : 0;
  uint4 A = ListRef16<L, NT>{w}.load();
  uint4 B = ListRef16<L, NT>{w + (size_t)min(1, blast) * BW}.load();
  A = make_uint4(nbody > 0 ? A.x : 0u, nbody > 0 ? A.y : 0u,
                 nbody > 0 ? A.z : 0u, nbody > 0 ? A.w : 0u);
  B = make_uint4(nbody > 0 ? B.x : 0u, nbody > 0 ? B.y : 0u,
                 nbody > 0 ? B.z : 0u, nbody > 0 ? B.w : 0u);
  ready(A);
  for (int h = 0; 2 * h < nbody; h++) {
    body(unpack16<L>(A),
         ListRef16<L, NT>{w + (size_t)min(2 * h + 2, blast) * BW}, A);
    if (L <= 2 || 2 * h + 1 < nbody) {
      body(unpack16<L>(B),
           ListRef16<L, NT>{w + (size_t)min(2 * h + 3, blast) * BW}, B);
      ready(A);
    }
  }
}

// 2F' with an atom's type (0 or 1) in the lowest mantissa bit, for the
// two-species force loop: of the FP32 values whose lowest bit is the type,
// the one nearest the FP64 value (v itself, or its neighbour on the FP64
// value's side: at most one unit in the last place, 1.2e-7 relative, and
// as often up as down).  The force loop reads the type back from the bit.
__device__ __forceinline__ float fold_type_bit(float v, double exact,
                                               int type) {
  const uint32_t b = __float_as_uint(v);
  if ((b & 1u) == (uint32_t)type) return v;
  if ((b & 0x7fffffffu) == 0u) return __uint_as_float(b + 1u);  // +-0
  // b + 1 is the neighbour away from zero, b - 1 the one towards zero
  const float up = __uint_as_float(b + 1u), dn = __uint_as_float(b - 1u);
  return fabs((double)up - exact) <= fabs((double)dn - exact) ? up : dn;
}

// Record j of an FP32 record array, addressed by a 32-bit byte offset from
// the array: the load takes the array's address from scalar registers and
// the offset from one vector register (no 64-bit address add per gather).
__device__ __forceinline__ float4 rec_at(const float4* __restrict__ rec,
                                         uint32_t j) {
  return *reinterpret_cast<const float4*>(
      reinterpret_cast<const char*>(rec) + j * 16u);
}

// Cell offset code and neighbour slot of a list entry in either format
// (16-bit: the base slot of the entry's code is the w of the wavefront's
// offset row, set per chunk); o receives the code's offset row.
template <bool W16>
__device__ __forceinline__ uint32_t entry_code(uint32_t e) {
  return W16 ? e >> IMDG_K16_BITS : e >> IMDG_IDX_BITS;
}
template <bool W16>
__device__ __forceinline__ int entry_slot(uint32_t e, const float4* off,
                                          float4& o) {
  o = off[entry_code<W16>(e)];
  return W16 ? __float_as_int(o.w) + (int)(e & IMDG_K16_MASK)
             : (int)(e & IMDG_IDX_MASK);
}

// The chunk's base slots into the w of the wavefront's offset row and its
// format (true: 16-bit entries).  Called by every lane of the wavefront.
__device__ __forceinline__ bool chunk_bases(const int* __restrict__ cmeta,
                                            int c, int lane, float4* off) {
  const int v = cmeta[(size_t)c * IMDG_META + min(lane, IMDG_META_WIDE)];
  if (lane < IMDG_NCODES) off[lane].w = __int_as_float(v);
  const int wide = __shfl(v, IMDG_META_WIDE, IMDG_WAVE);
  wave_lds_sync();
  return wide == 0;
}

// One walk over a slot's list in the chunk's format; NT: the list words'
// loads are non-temporal (the list_nt policy, chosen per launch).
template <int L, bool NT, typename F>
__device__ __forceinline__ void walk_either(bool w16, const uint32_t* nbr,
                                            int i, int kmax, int s, int nnb,
                                            F body) {
  if (w16)
    list_walk16<L, NT>(list_words16<L>(nbr, i, kmax, s), nnb, body);
  else
    list_walk<L, NT>(list_words<L>(nbr, i, kmax, s), nnb, body);
}

// Single-species column of the density loop.
template <typename R>
struct DensityCol {
  int kb, len;
  R begin, end, invstep;
};

// Single-species columns of the force loop (pair, rho).
template <typename R>
struct ForceCols {
  int pk, plen;
  R pb, pe, pis, p2is;
  int rk, rlen;
  R rb, re, ris, r2is;
};

// Table argument of the FP32 segment tables (one grid): t = (r^2 - begin)
// istep as fma(r^2, istep, t0) clamped to [0, len - 1/2] (med3), segment
// k = (int)t, fraction chi = fract(t) + fma(r^2, ilo, t0lo): the low parts
// of the split constants (TabDev::ilo, t0lo) go into the fraction, where
// FP32 still resolves them, instead of shifting every pair of a shell the
// same way.
struct TArg {
  float istep, ilo, t0, t0lo, tz;
};

template <typename R>
__device__ __forceinline__ TArg targ(const TabDev<R>& t) {
  TArg a;
  a.istep = (float)t.invstep[0];
  a.ilo = (float)t.ilo[0];
  a.t0 = (float)t.t0[0];
  a.t0lo = (float)t.t0lo[0];
  a.tz = (float)t.len[0] - 0.5f;
  return a;
}

__device__ __forceinline__ int tseg(float r2, const TArg& A, float& chi) {
  float t = __builtin_amdgcn_fmed3f(fmaf(r2, A.istep, A.t0), 0.0f, A.tz);
  chi = __builtin_amdgcn_fractf(t) + fmaf(r2, A.ilo, A.t0lo);
  return (int)t;
}

// Four neighbours of the density loop: records and offsets, the prefetch,
// r^2 and segments, all knot loads together, accumulation with mask
// multiplies (no branches around the loads).
template <typename R, bool NT1, bool CUB, bool EE, typename R4, typename R2,
          typename PF>
__device__ __forceinline__ void density_group(
    uint4 e4, PF pf, uint4& slot, const R4* __restrict__ rec,
    const R4* off, const R4& ui, int it, int nt, const R2* __restrict__ K,
    const ColPar<R>* cpr, const DensityCol<R>& c0, R& rho, R& pee,
    int& is_short, R& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  R4 uj[4], o[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    if (NT1) {
      // one species: the type is not needed, gather 12 B instead of 16
      const typename V2<R>::T* p2 =
          reinterpret_cast<const typename V2<R>::T*>(rec +
                                                     (en[u] & IMDG_IDX_MASK));
      typename V2<R>::T xy = p2[0];
      uj[u].x = xy.x;
      uj[u].y = xy.y;
      uj[u].z = reinterpret_cast<const R*>(p2)[2];
      uj[u].w = (R)0;
    } else {
      uj[u] = rec[en[u] & IMDG_IDX_MASK];
    }
    o[u] = off[en[u] >> IMDG_IDX_BITS];
  }
  // issued after the gathers: loads complete in order, and the gathers
  // must not wait for this (HBM) one
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
  R r2[4], chi[4], cb[4], ce[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    R dx = (uj[u].x - ui.x) + o[u].x;
    R dy = (uj[u].y - ui.y) + o[u].y;
    R dz = (uj[u].z - ui.z) + o[u].z;
    r2[u] = dx * dx + dy * dy + dz * dz;
    if (NT1) {
      kk[u] = c0.kb + tab_knot<CUB>(seg_locate(r2[u], c0.begin, c0.end,
                                               c0.invstep, c0.len, chi[u]));
      cb[u] = c0.begin;
      ce[u] = c0.end;
    } else {
      ColPar<R> cp = cpr[it * nt + (int)uj[u].w];
      kk[u] = cp.off + tab_knot<CUB>(seg_locate(r2[u], cp.begin, cp.end,
                                                cp.invstep, cp.len, chi[u]));
      cb[u] = cp.begin;
      ce[u] = cp.end;
    }
  }
  R2 p[4], q[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    p[u] = K[kk[u]];
    q[u] = K[kk[u] + 1];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    bool in = r2[u] < ce[u];
    R v = tab_val<CUB, R>(p[u], q[u], chi[u]);
    rho = fma(in ? (R)1 : (R)0, v, rho);
    // EEAM: P += rho^2 (src/imd_forces_nbl.c:591-593)
    if (EE) pee = fma(in ? v : (R)0, v, pee);
    if (NT1)
      r2min = r2[u] < r2min ? r2[u] : r2min;  // r2 < begin implies in range
    else
      is_short |= (in && r2[u] < cb[u]);
  }
}

// FP32 single-species variant of density_group on the rho segment
// polynomials DQ[3k .. 3k+2] = (p0, A, B) (three packed floats: 12 B of LDS
// per segment instead of 16): value p0 + chi (A + B chi); the table
// argument is clamped to [0, len - 1/2] (med3), so r^2 at or beyond end
// reads the zero segment len - 1 and no cutoff test is needed.  r^2 keeps
// the offset remainder (sep_r2): without it every rho shifts by a uniform
// ~1e-7 (relative), and through F' the virial by ~1e-4.  Records are
// addressed by a 32-bit byte offset from the array (scalar base + vector
// offset, no 64-bit address add).
template <typename PF>
__device__ __forceinline__ void density_group_poly(
    uint4 e4, PF pf, uint4& slot, const float4* __restrict__ rec,
    const float4* off, const float4* olo, const float4& ui,
    const float* __restrict__ DQ, const TArg& ta, float& rho,
    float& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  float ux[4], uy[4], uz[4];
  float4 o[4], ol[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    // the type is not needed: gather 12 B
    const uint32_t j = (uint32_t)entry_slot<PF::W16>(en[u], off, o[u]);
    const float* p = reinterpret_cast<const float*>(
        reinterpret_cast<const char*>(rec) + j * 16u);
    float2 xy = *reinterpret_cast<const float2*>(p);
    ux[u] = xy.x;
    uy[u] = xy.y;
    uz[u] = p[2];
  }
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
#pragma unroll
  for (int u = 0; u < 4; u++) ol[u] = olo[entry_code<PF::W16>(en[u])];
  float r2[4], chi[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float dx = (ux[u] - ui.x) + o[u].x;
    float dy = (uy[u] - ui.y) + o[u].y;
    float dz = (uz[u] - ui.z) + o[u].z;
    r2[u] = sep_r2(dx, dy, dz, ol[u]);
    kk[u] = tseg(r2[u], ta, chi[u]);
  }
  float q0[4], q1[4], q2[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    q0[u] = DQ[3 * kk[u]];
    q1[u] = DQ[3 * kk[u] + 1];
    q2[u] = DQ[3 * kk[u] + 2];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    rho += fmaf(chi[u], fmaf(chi[u], q2[u], q1[u]), q0[u]);
    r2min = fminf(r2min, r2[u]);
  }
}

// Four neighbours of the force loop (same phases as density_group).
template <typename R, bool NT1, bool HP, bool HE, bool CUB, bool EE,
          typename R4, typename R2, typename PF>
__device__ __forceinline__ void force_group(
    uint4 e4, PF pf, uint4& slot, const R4* __restrict__ rec,
    const signed char* __restrict__ atype, const R4* off, const R4& ui, R dFi,
    int it, int nt, const R2* __restrict__ K, const ColPar<R>* cpp,
    const ColPar<R>* cpr, const ForceCols<R>& c, const R* __restrict__ eedm,
    R dmi, R& fx, R& fy, R& fz, R& ep, R& vir, int& is_short) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  R4 uj[4], o[4];
  int jt[4];
  R dmj[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    uj[u] = rec[en[u] & IMDG_IDX_MASK];
    jt[u] = NT1 ? 0 : (int)atype[en[u] & IMDG_IDX_MASK];
    dmj[u] = EE ? eedm[en[u] & IMDG_IDX_MASK] : (R)0;
    o[u] = off[en[u] >> IMDG_IDX_BITS];
  }
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
  R dx[4], dy[4], dz[4], r2[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    dx[u] = (uj[u].x - ui.x) + o[u].x;
    dy[u] = (uj[u].y - ui.y) + o[u].y;
    dz[u] = (uj[u].z - ui.z) + o[u].z;
    r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
  }
  R gg[4];
  if (NT1) {
    R cp[4], cr[4];
    int kp[4], kr[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (HP)
        kp[u] = c.pk + tab_knot<CUB>(
                           seg_locate(r2[u], c.pb, c.pe, c.pis, c.plen, cp[u]));
      if (HE)
        kr[u] = c.rk + tab_knot<CUB>(
                           seg_locate(r2[u], c.rb, c.re, c.ris, c.rlen, cr[u]));
    }
    R2 p0[4], p1[4], q0[4], q1[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (HP) {
        p0[u] = K[kp[u]];
        p1[u] = K[kp[u] + 1];
      }
      if (HE) {
        q0[u] = K[kr[u]];
        q1[u] = K[kr[u] + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      R g1 = (R)0;
      if (HP) {
        bool in = r2[u] <= c.pe;
        R m = in ? (R)1 : (R)0;
        g1 = m * tab_grad<CUB, R>(p0[u], p1[u], cp[u], c.p2is);
        ep = fma(m, tab_val<CUB, R>(p0[u], p1[u], cp[u]), ep);
        is_short |= (in && r2[u] < c.pb);
      }
      if (HE) {
        bool in = r2[u] < c.re;
        R m = in ? (R)0.5 : (R)0;
        R rg = tab_grad<CUB, R>(q0[u], q1[u], cr[u], c.r2is);
        g1 = fma(m * (dFi + uj[u].w), rg, g1);
        if (EE) {
          // (DM_i + DM_j) rho rho' (one species: rho_i = rho_j)
          R rv = tab_val<CUB, R>(q0[u], q1[u], cr[u]);
          g1 = fma((in ? dmi + dmj[u] : (R)0) * rv, rg, g1);
        }
        is_short |= (in && r2[u] < c.rb);
      }
      gg[u] = g1;
    }
  } else {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int col = it * nt + jt[u], colT = jt[u] * nt + it;
      R g1 = (R)0;
      if (HP) {
        ColPar<R> cp = cpp[col];
        if (r2[u] <= cp.end) {
          R chi;
          int k = tab_knot<CUB>(
              seg_locate(r2[u], cp.begin, cp.end, cp.invstep, cp.len, chi));
          R2 p = K[cp.off + k], q = K[cp.off + k + 1];
          g1 += tab_grad<CUB, R>(p, q, chi, cp.twoistep);
          ep += tab_val<CUB, R>(p, q, chi);
          is_short |= (r2[u] < cp.begin);
        }
      }
      if (HE) {
        // rho'_i from column jt*nt+it, rho'_j from it*nt+jt
        // (src/imd_forces_nbl.c:1156-1203)
        ColPar<R> ci = cpr[colT], cj = cpr[col];
        if (r2[u] < ci.end || r2[u] < cj.end) {
          R chi;
          int k = tab_knot<CUB>(
              seg_locate(r2[u], ci.begin, ci.end, ci.invstep, ci.len, chi));
          is_short |= (r2[u] < ci.begin);
          R rpi = tab_grad<CUB, R>(K[ci.off + k], K[ci.off + k + 1], chi,
                                   ci.twoistep);
          R rvi = EE ? tab_val<CUB, R>(K[ci.off + k], K[ci.off + k + 1], chi)
                     : (R)0;
          R rpj = rpi, rvj = rvi;
          if (colT != col) {
            k = tab_knot<CUB>(
                seg_locate(r2[u], cj.begin, cj.end, cj.invstep, cj.len, chi));
            is_short |= (r2[u] < cj.begin);
            rpj = tab_grad<CUB, R>(K[cj.off + k], K[cj.off + k + 1], chi,
                                   cj.twoistep);
            if (EE)
              rvj = tab_val<CUB, R>(K[cj.off + k], K[cj.off + k + 1], chi);
          }
          g1 += (R)0.5 * (dFi * rpj + uj[u].w * rpi);
          // EEAM: DM_i rho_j rho'_j + DM_j rho_i rho'_i
          // (src/imd_forces_nbl.c:1204-1207)
          if (EE) g1 += dmi * rvj * rpj + dmj[u] * rvi * rpi;
        }
      }
      gg[u] = g1;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    fx = fma(dx[u], gg[u], fx);
    fy = fma(dy[u], gg[u], fy);
    fz = fma(dz[u], gg[u], fz);
    vir = fma(r2[u], gg[u], vir);
  }
}

// Four neighbours of the single-species force loop when the pair and rho
// tables share their grid: one segment per neighbour, its record S[k] =
// (phi p0, phi dv, phi d2v, rho dv) and Sr[k] = rho d2v.  The common factor
// 2 istep of both gradients is applied once per atom by the caller (sums
// s* are in units of 1 / (2 istep)); hFi = F'(rho_i) = 2F'/2.
template <typename R, typename R4, typename PF>
__device__ __forceinline__ void force_group_sg(
    uint4 e4, PF pf, uint4& slot, const R4* __restrict__ rec,
    const R4* off, const R4& ui, R hFi, const R4* __restrict__ S,
    const R* __restrict__ Sr, int len, R begin, R end, R invstep, R& sx,
    R& sy, R& sz, R& ep, R& svir, R& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  R4 uj[4], o[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    uj[u] = rec[en[u] & IMDG_IDX_MASK];
    o[u] = off[en[u] >> IMDG_IDX_BITS];
  }
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
  R dx[4], dy[4], dz[4], r2[4], chi[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    dx[u] = (uj[u].x - ui.x) + o[u].x;
    dy[u] = (uj[u].y - ui.y) + o[u].y;
    dz[u] = (uj[u].z - ui.z) + o[u].z;
    r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
    kk[u] = seg_locate(r2[u], begin, end, invstep, len, chi[u]);
  }
  R4 s4[4];
  R d2r[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    s4[u] = S[kk[u]];
    d2r[u] = Sr[kk[u]];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    R c = chi[u] - (R)0.5;
    R gp = fma(c, s4[u].z, s4[u].y);
    R gr = fma(c, d2r[u], s4[u].w);
    R vp = fma(chi[u], fma((R)0.5 * (chi[u] - (R)1), s4[u].z, s4[u].y),
               s4[u].x);
    // pair: r2 <= end (PAIR_INT2 guard); rho: r2 < end (VAL/DERIV_FUNC2)
    bool inp = r2[u] <= end, inr = r2[u] < end;
    R h = fma((R)0.5, uj[u].w, hFi);
    R g = fma(h, inr ? gr : (R)0, inp ? gp : (R)0);
    ep += inp ? vp : (R)0;
    sx = fma(dx[u], g, sx);
    sy = fma(dy[u], g, sy);
    sz = fma(dz[u], g, sz);
    svir = fma(r2[u], g, svir);
    r2min = r2[u] < r2min ? r2[u] : r2min;
  }
}

// FP32 variant of force_group_sg on the polynomial segment table: Q[k] =
// (phi p0, phi A, phi B, phi A_lo), Rr[k] = (rho A, rho D); value p0 + chi
// (A + B chi), gradients / (2 istep) A + B (2 chi) (pair; 2B is phi's D
// exactly) and A + D chi (rho).  The table argument is clamped to [0, len -
// 1/2] (tseg): r^2 below begin reads segment 0 at chi = 0 like the
// reference, r^2 at or beyond end reads the zero segment len - 1, so no
// cutoff test is needed.  svlo sums r^2 A_lo, the part of the pair
// gradient that FP32 drops from A, for the virial.
template <typename PF>
__device__ __forceinline__ void force_group_sgf(
    uint4 e4, PF pf, uint4& slot, const float4* __restrict__ rec,
    const float4* off, const float4* olo, const float4& ui, float hFi,
    const float4* __restrict__ Q, const float2* __restrict__ Rr,
    const TArg& ta, float& sx, float& sy, float& sz, float& ep, float& svir,
    float& svlo) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  float4 uj[4], o[4], ol[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    uj[u] = rec_at(rec, (uint32_t)entry_slot<PF::W16>(en[u], off, o[u]));
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
#pragma unroll
  for (int u = 0; u < 4; u++) ol[u] = olo[entry_code<PF::W16>(en[u])];
  float dx[4], dy[4], dz[4], r2[4], chi[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    dx[u] = (uj[u].x - ui.x) + o[u].x;
    dy[u] = (uj[u].y - ui.y) + o[u].y;
    dz[u] = (uj[u].z - ui.z) + o[u].z;
    r2[u] = sep_r2(dx[u], dy[u], dz[u], ol[u]);
    kk[u] = tseg(r2[u], ta, chi[u]);
  }
  float4 q[4];
  float2 rr[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    q[u] = Q[kk[u]];
    rr[u] = Rr[kk[u]];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float gp = fmaf(chi[u] + chi[u], q[u].z, q[u].y);
    float gr = fmaf(chi[u], rr[u].y, rr[u].x);
    float vp = fmaf(chi[u], fmaf(chi[u], q[u].z, q[u].y), q[u].x);
    float h = fmaf(0.5f, uj[u].w, hFi);
    float g = fmaf(h, gr, gp);
    ep += vp;
    sx = fmaf(dx[u], g, sx);
    sy = fmaf(dy[u], g, sy);
    sz = fmaf(dz[u], g, sz);
    svir = fmaf(r2[u], g, svir);
    svlo = fmaf(r2[u], q[u].w, svlo);
  }
}

// Several species on one grid (FP32): the record pair S[2 (p len + k)] of
// type pair p = it*nt + jt is (phi p0, phi A, phi B, phi A_lo) of pair
// column p and (A, D) of rho columns p and jt*nt + it; gradients in units
// of 1/(2 istep) and the virial's A_lo sum as in force_group_sgf.  rho'_j
// (column p) goes with 2F'_i, rho'_i (column jt*nt + it) with 2F'_j
// (src/imd_forces_nbl.c:1156-1203).
//
// LR: the rho parts come from the per-type table FR in LDS instead (rho
// column it*nt + jt the same for every it): FR[t len + k] = (A, D) of the
// density due to a t-atom, the same FP32 values as the records' second
// half, so the results do not change; one gather per neighbour less.  S
// then holds only the pair halves, S[p len + k].
template <bool TB, bool LR, typename PF>
__device__ __forceinline__ void force_group_sgn(
    uint4 e4, PF pf, uint4& slot, const float4* __restrict__ rec,
    const signed char* __restrict__ atype, const float4* off,
    const float4* olo, const float4& ui, float hFi,
    const float4* __restrict__ S, const float2* FR, int it, int pbase,
    int len, const TArg& ta, float& sx, float& sy, float& sz, float& ep,
    float& svir, float& svlo, float& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  float4 uj[4], o[4], ol[4];
  int jt[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int j = entry_slot<PF::W16>(en[u], off, o[u]);
    uj[u] = rec_at(rec, (uint32_t)j);
    if (!TB) jt[u] = atype[j];
  }
  if (TB) {
#pragma unroll
    for (int u = 0; u < 4; u++) jt[u] = (int)(__float_as_uint(uj[u].w) & 1u);
  }
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
#pragma unroll
  for (int u = 0; u < 4; u++) ol[u] = olo[entry_code<PF::W16>(en[u])];
  float dx[4], dy[4], dz[4], r2[4], chi[4];
  int kk[4], sg[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    dx[u] = (uj[u].x - ui.x) + o[u].x;
    dy[u] = (uj[u].y - ui.y) + o[u].y;
    dz[u] = (uj[u].z - ui.z) + o[u].z;
    r2[u] = sep_r2(dx[u], dy[u], dz[u], ol[u]);
    sg[u] = tseg(r2[u], ta, chi[u]);
    kk[u] = (LR ? 1 : 2) * ((pbase + jt[u]) * len + sg[u]);
  }
  float4 q[4], rr[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    q[u] = S[kk[u]];
    if (LR) {
      const float2 fj = FR[jt[u] * len + sg[u]], fi = FR[it * len + sg[u]];
      rr[u] = make_float4(fj.x, fj.y, fi.x, fi.y);
    } else {
      rr[u] = S[kk[u] + 1];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float gp = fmaf(chi[u] + chi[u], q[u].z, q[u].y);
    float grj = fmaf(chi[u], rr[u].y, rr[u].x);
    float gri = fmaf(chi[u], rr[u].w, rr[u].z);
    float vp = fmaf(chi[u], fmaf(chi[u], q[u].z, q[u].y), q[u].x);
    float g = fmaf(hFi, grj, fmaf(0.5f * uj[u].w, gri, gp));
    ep += vp;
    sx = fmaf(dx[u], g, sx);
    sy = fmaf(dy[u], g, sy);
    sz = fmaf(dz[u], g, sz);
    svir = fmaf(r2[u], g, svir);
    svlo = fmaf(r2[u], q[u].w, svlo);
    r2min = fminf(r2min, r2[u]);
  }
}

// Several species on one grid (FP32): density records DQ[p len + k] = (rho
// p0, A, B, 0) of column p = it*nt + jt, jt from the gathered record.
template <typename PF>
__device__ __forceinline__ void density_group_polyn(
    uint4 e4, PF pf, uint4& slot, const float4* __restrict__ rec,
    const float4* off, const float4* olo, const float4& ui,
    const float4* __restrict__ DQ, int pbase, int len, const TArg& ta,
    float& rho, float& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  float4 uj[4], o[4], ol[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    uj[u] = rec_at(rec, (uint32_t)entry_slot<PF::W16>(en[u], off, o[u]));
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
#pragma unroll
  for (int u = 0; u < 4; u++) ol[u] = olo[entry_code<PF::W16>(en[u])];
  float r2[4], chi[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float dx = (uj[u].x - ui.x) + o[u].x;
    float dy = (uj[u].y - ui.y) + o[u].y;
    float dz = (uj[u].z - ui.z) + o[u].z;
    r2[u] = sep_r2(dx, dy, dz, ol[u]);
    kk[u] = (pbase + (int)uj[u].w) * len + tseg(r2[u], ta, chi[u]);
  }
  float4 q[4];
#pragma unroll
  for (int u = 0; u < 4; u++) q[u] = DQ[kk[u]];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    rho += fmaf(chi[u], fmaf(chi[u], q[u].z, q[u].y), q[u].x);
    r2min = fminf(r2min, r2[u]);
  }
}

// Several species on one grid, rho column it*nt + jt the same for every
// it: DR[3 (jt len + k) + 0..2] = (rho p0, A, B) of the density due to a
// jt-atom, staged in LDS (the same FP32 values as the records of
// density_group_polyn, so the results do not change).
template <typename PF>
__device__ __forceinline__ void density_group_polyl(
    uint4 e4, PF pf, uint4& slot, const float4* __restrict__ rec,
    const float4* off, const float4* olo, const float4& ui, const float* DR,
    int len, const TArg& ta, float& rho, float& r2min) {
  uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
  float4 uj[4], o[4], ol[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
    uj[u] = rec_at(rec, (uint32_t)entry_slot<PF::W16>(en[u], off, o[u]));
  __builtin_amdgcn_sched_barrier(0);
  slot = pf.load();
#pragma unroll
  for (int u = 0; u < 4; u++) ol[u] = olo[entry_code<PF::W16>(en[u])];
  float r2[4], chi[4];
  int kk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    float dx = (uj[u].x - ui.x) + o[u].x;
    float dy = (uj[u].y - ui.y) + o[u].y;
    float dz = (uj[u].z - ui.z) + o[u].z;
    r2[u] = sep_r2(dx, dy, dz, ol[u]);
    kk[u] = 3 * ((int)uj[u].w * len + tseg(r2[u], ta, chi[u]));
  }
  float q0[4], q1[4], q2[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    q0[u] = DR[kk[u]];
    q1[u] = DR[kk[u] + 1];
    q2[u] = DR[kk[u] + 2];
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    rho += fmaf(chi[u], fmaf(chi[u], q2[u], q1[u]), q0[u]);
    r2min = fminf(r2min, r2[u]);
  }
}

// ---------------------------------------------------------------------------
// density loop
// ---------------------------------------------------------------------------

// Host density, embedding energy and 2F'(rho) (src/imd_forces_nbl.c:586-611,
// 1079-1112): rho_i = sum_j rho_{type(i) type(j)}(r_ij^2) over the full
// list, then F and 2F' from the embedding column of type(i).  Writes the
// (u, 2F') records the force loop gathers.  L lanes per slot; CUB: the
// tables are cubic segments (4-point or spline interpolation); NTL: the
// FP32 one-grid walks load the list words non-temporally.
template <typename R, bool NT1, bool LT, int L, bool CUB, bool EE,
          bool NTL = false>
__global__ void __launch_bounds__(IMDG_STEP_THREADS)
    k_density(StepArgs<R> a, const TabDev<R>* __restrict__ tabs) {
  typedef typename V4<R>::T R4;
  typedef typename V2<R>::T R2;
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  // per wavefront: the offset row, then its remainder row (one address
  // register serves both: the remainder is at a constant distance)
  __shared__ R4 wofs[IMDG_STEP_WAVES][2][IMDG_NCODES];
  __shared__ ColPar<R> cpr[IMDG_MAXCOLS];
  if (a.ctl->frozen) return;
  const TabDev<R>& tr = tabs[1];
  const TabDev<R>& te = tabs[2];
  const int klo = LT ? a.lds_lo : 0;
  if (LT) stage_knots(a.knots, a.lds_lo, a.lds_n, (R2*)dyn_lds);
  if (!NT1) stage_colpar(tr, klo, cpr);
  __syncthreads();
  const R2* __restrict__ K = LT ? (const R2*)dyn_lds : a.knots;
  const R4* __restrict__ rec = a.rec_a;
  const int nt = a.ntypes;
  // single-species column
  const int kb0 = tr.off[0] - klo, len0 = tr.len[0];
  const R b0 = tr.begin[0], e0 = tr.end[0], is0 = tr.invstep[0];

  Sched S = make_sched(a.nchunk, a.xcd_remap);
  const int lane = threadIdx.x & (IMDG_WAVE - 1);
  const int ga = lane / L, gs = lane % L;
  R4* off = wofs[threadIdx.x / IMDG_WAVE][0];
  R4* olo = wofs[threadIdx.x / IMDG_WAVE][1];
  int cur = -1, is_short = 0;
  double acc = 0.0;
  for (int c = S.c; c < S.c1; c += S.stride) {
    int r = a.blk_rep[c / (IMDG_BLOCK / IMDG_CHUNK)];
    const RepDev& Rp = a.rep[r];
    if (r != cur) {
      wave_offsets(Rp, off, olo);
      cur = r;
    }
    bool w16 = false;
    if constexpr (sizeof(R) == 4)
      if (a.list16) w16 = chunk_bases(a.cmeta, c, lane, off);
    R rho_keep = (R)0, pee_keep = (R)0;
#pragma unroll 1
    for (int p = 0; p < L; p++) {
      int i = c * IMDG_CHUNK + p * (IMDG_WAVE / L) + ga;
      R rho = (R)0, pee = (R)0;
      if (i - Rp.atom_off < Rp.natoms) {
        R4 ui = rec[i];
        int it = NT1 ? 0 : (int)ui.w;
        DensityCol<R> col0 = {kb0, len0, b0, e0, is0};
        R r2min = (R)3.0e38;
        const uint32_t* nw = list_words<L>(a.nbr, i, a.kmax, gs);
        if constexpr (sizeof(R) == 4 && NT1 && !CUB && !EE) {
          const float* __restrict__ DQ =
              reinterpret_cast<const float*>(K + (a.dpoly - klo));
          const TArg ta = targ(tr);
          walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                         [&](uint4 e4, auto pf, uint4& slot) {
                           density_group_poly(e4, pf, slot, rec, off, olo,
                                              ui, DQ, ta, rho, r2min);
                         });
        } else {
          auto generic = [&]() {
            list_walk<L>(nw, a.nnb[i],
                         [&](uint4 e4, ListRef<L> pf, uint4& slot) {
                           density_group<R, NT1, CUB, EE>(
                               e4, pf, slot, rec, off, ui, it, nt, K, cpr,
                               col0, rho, pee, is_short, r2min);
                         });
          };
          if constexpr (sizeof(R) == 4 && !NT1 && !CUB && !EE && LT) {
            if (a.drho >= 0) {
              // several species on one grid, per-type rho segments in LDS
              const float* DR =
                  reinterpret_cast<const float*>(K + (a.drho - klo));
              const TArg ta = targ(tr);
              walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                             [&](uint4 e4, auto pf, uint4& slot) {
                               density_group_polyl(e4, pf, slot, rec, off,
                                                   olo, ui, DR, len0, ta, rho,
                                                   r2min);
                             });
              is_short |= (r2min < b0);
            } else {
              generic();
            }
          } else if constexpr (sizeof(R) == 4 && !NT1 && !CUB && !EE) {
            if (a.dpolyn >= 0) {
              // several species on one grid: segment records in HBM/L2
              const float4* __restrict__ DQ =
                  reinterpret_cast<const float4*>(a.knots + a.dpolyn);
              const TArg ta = targ(tr);
              walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                             [&](uint4 e4, auto pf, uint4& slot) {
                               density_group_polyn(e4, pf, slot, rec, off,
                                                   olo, ui, DQ, it * nt, len0,
                                                   ta, rho, r2min);
                             });
              is_short |= (r2min < b0);
            } else {
              generic();
            }
          } else {
            generic();
          }
        }
        if (NT1) is_short |= (r2min < b0);
      }
      rho = lane_group_sum<L>(rho);
      if (gs == p) rho_keep = rho;
      if (EE) {
        pee = lane_group_sum<L>(pee);
        if (gs == p) pee_keep = pee;
      }
    }
    rho_keep = to_slot_lane<L>(rho_keep, lane);
    if (EE) pee_keep = to_slot_lane<L>(pee_keep, lane);
    int i = c * IMDG_CHUNK + lane;
    double emb = 0.0;
    if (i - Rp.atom_off < Rp.natoms) {
      R4 ui = rec[i];
      int it = NT1 ? 0 : (int)ui.w;
      R rho = rho_keep;
      // embedding: PAIR_INT on the embedding column of the type, argument
      // rho (knots in HBM; one lookup per atom), in FP64 (the FP32 kernels
      // read FP64 knots of their own)
      double Fd, dFd;
      if constexpr (sizeof(R) == 4) {
        const Tab64& E = a.tab64[0];
        const double2* __restrict__ ke = a.knots64 + E.off[it];
        double chi;
        int k = tab_knot<CUB>(seg_locate((double)rho, E.begin[it], E.end[it],
                                         E.invstep[it], E.len[it], chi));
        double2 p = ke[k], q = ke[k + 1];
        Fd = tab_val<CUB, double>(p, q, chi);
        dFd = tab_grad<CUB, double>(p, q, chi, E.twoistep[it]);
      } else {
        const R2* __restrict__ ke = a.knots + te.off[it];
        R chi;
        int k = tab_knot<CUB>(seg_locate(rho, te.begin[it], te.end[it],
                                         te.invstep[it], te.len[it], chi));
        R2 p = ke[k], q = ke[k + 1];
        Fd = tab_val<CUB, R>(p, q, chi);
        dFd = tab_grad<CUB, R>(p, q, chi, te.twoistep[it]);
      }
      if (EE) {
        // EEAM energy term E(P) and EAM_DM = 2E'(P)
        // (src/imd_forces_nbl.c:1090-1095)
        double Em, dEm;
        if constexpr (sizeof(R) == 4) {
          const Tab64& E = a.tab64[1];
          const double2* __restrict__ km = a.knots64 + E.off[it];
          double chm;
          int kq = tab_knot<CUB>(seg_locate((double)pee_keep, E.begin[it],
                                            E.end[it], E.invstep[it],
                                            E.len[it], chm));
          double2 pm = km[kq], qm = km[kq + 1];
          Em = tab_val<CUB, double>(pm, qm, chm);
          dEm = tab_grad<CUB, double>(pm, qm, chm, E.twoistep[it]);
        } else {
          const TabDev<R>& tm = tabs[3];
          const R2* __restrict__ km = a.knots + tm.off[it];
          R chm;
          int kq = tab_knot<CUB>(seg_locate(pee_keep, tm.begin[it],
                                            tm.end[it], tm.invstep[it],
                                            tm.len[it], chm));
          R2 pm = km[kq], qm = km[kq + 1];
          Em = tab_val<CUB, R>(pm, qm, chm);
          dEm = tab_grad<CUB, R>(pm, qm, chm, tm.twoistep[it]);
        }
        Fd += Em;
        a.eedm[i] = (R)dEm;
        if (a.store) a.eeP[i] = pee_keep;
      }
      const R F = (R)Fd, dF = (R)dFd;
      R4 rb;
      rb.x = ui.x;
      rb.y = ui.y;
      rb.z = ui.z;
      rb.w = dF;
      if constexpr (sizeof(R) == 4)
        if (a.type_bit) rb.w = fold_type_bit(dF, dFd, it);
      a.rec_b[i] = rb;
      if (a.store) {
        a.rho[i] = rho;
        a.epot[i] = F;
      }
      emb = Fd;
    }
    if (a.part_chunk) {
      double s = wave_sum(emb);
      if (lane == 0) a.partA[c] = s;
    } else {
      acc += emb;
    }
  }
  if (!a.part_chunk) {
    double s = wave_sum(acc);
    block_partial<1>(a.partA + blockIdx.x, &s, -1);
  }
  flag_short(a.ctl, is_short);
}

// ---------------------------------------------------------------------------
// force loop (+ integration)
// ---------------------------------------------------------------------------

// Pair + EAM forces (src/imd_forces_nbl.c:490-584, 1117-1322) as a gather
// over the full list of (u, 2F') records; with INTEGRATE the leapfrog step,
// the skin test and the step's reductions follow in the same thread
// (calc_forces, move_atoms and check_nblist of one main_loop iteration in
// one launch).  L lanes per slot; NTL: the FP32 one-grid walks load the
// list words non-temporally.
template <typename R, int MODE, bool INTEGRATE, bool HP, bool HE, bool LT,
          int L, bool CUB, bool EE, bool NTL = false>
__global__ void __launch_bounds__((ForceThreads<R, MODE>::value))
    k_force(StepArgs<R> a, const TabDev<R>* __restrict__ tabs) {
  constexpr bool NT1 = MODE != FM_NTN && MODE != FM_SGN;
  typedef typename V4<R>::T R4;
  typedef typename V2<R>::T R2;
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  // per wavefront: the offset row, then its remainder row (one address
  // register serves both: the remainder is at a constant distance)
  __shared__ R4 wofs[IMDG_STEP_WAVES][2][IMDG_NCODES];
  __shared__ ColPar<R> cpp[IMDG_MAXCOLS], cpr[IMDG_MAXCOLS];
  if (a.ctl->frozen) return;
  const TabDev<R>& tp = tabs[0];
  const TabDev<R>& tr = tabs[1];
  const int klo = LT ? a.lds_lo : 0;
  if (LT) stage_knots(a.knots, a.lds_lo, a.lds_n, (R2*)dyn_lds);
  if (!NT1) {
    if (HP) stage_colpar(tp, klo, cpp);
    if (HE) stage_colpar(tr, klo, cpr);
  }
  __syncthreads();
  const R2* __restrict__ K = LT ? (const R2*)dyn_lds : a.knots;
  // the gathered records (rec_b): (u, 2F') from the density loop with EAM,
  // a copy of (u, type) without; the integration below rewrites rec_a
  const R4* __restrict__ rec = a.rec_b;
  const signed char* __restrict__ atype = a.atype;
  const int nt = a.ntypes;
  // single-species columns
  const int pk0 = HP ? tp.off[0] - klo : 0, plen = HP ? tp.len[0] : 0;
  const R pb = HP ? tp.begin[0] : (R)0, pe = HP ? tp.end[0] : (R)0,
          pis = HP ? tp.invstep[0] : (R)0, p2is = HP ? tp.twoistep[0] : (R)0;
  const int rk0 = HE ? tr.off[0] - klo : 0, rlen = HE ? tr.len[0] : 0;
  const R rb = HE ? tr.begin[0] : (R)0, re = HE ? tr.end[0] : (R)0,
          ris = HE ? tr.invstep[0] : (R)0, r2is = HE ? tr.twoistep[0] : (R)0;

  Sched S = make_sched(a.nchunk, a.xcd_remap);
  const int lane = threadIdx.x & (IMDG_WAVE - 1);
  const int ga = lane / L, gs = lane % L;
  R4* off = wofs[threadIdx.x / IMDG_WAVE][0];
  R4* olo = wofs[threadIdx.x / IMDG_WAVE][1];
  int cur = -1, is_short = 0;
  double acc[PB_N] = {0, 0, 0, 0, 0, 0};
  for (int c = S.c; c < S.c1; c += S.stride) {
    int r = a.blk_rep[c / (IMDG_BLOCK / IMDG_CHUNK)];
    const RepDev& Rp = a.rep[r];
    if (r != cur) {
      wave_offsets(Rp, off, olo);
      cur = r;
    }
    bool w16 = false;
    if constexpr (sizeof(R) == 4)
      if (a.list16) w16 = chunk_bases(a.cmeta, c, lane, off);
    // the slot's sums kept by this lane (pass gs); vlo: the virial's sum
    // of the pair gradients' FP32 remainders (FP32 one-grid loops)
    R kx = 0, ky = 0, kz = 0, kep = 0, kvir = 0, kvlo = 0;
#pragma unroll 1
    for (int p = 0; p < L; p++) {
      int i = c * IMDG_CHUNK + p * (IMDG_WAVE / L) + ga;
      R fx = 0, fy = 0, fz = 0, ep = 0, vir = 0, vlo = 0;
      if (i - Rp.atom_off < Rp.natoms) {
        R4 ui = rec[i];
        int it = NT1 ? 0 : (int)atype[i];
        R dFi = HE ? ui.w : (R)0;
        const uint32_t* nw = list_words<L>(a.nbr, i, a.kmax, gs);
        if constexpr (MODE == FM_SGN && sizeof(R) == 4 && !CUB && !EE) {
          // several species on one grid: sums in units of 1/(2 istep)
          R r2min = (R)3.0e38;
          // LT: the per-type rho segments are in LDS (the launch stages
          // nothing else for this loop), the pair halves of the records
          // in their own table
          const float4* __restrict__ S = reinterpret_cast<const float4*>(
              a.knots + (LT ? a.pairn : a.jointn));
          const float2* FR =
              LT ? reinterpret_cast<const float2*>(K + (a.frho - klo))
                 : nullptr;
          const TArg ta = targ(tp);
          if (a.type_bit)
            walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                           [&](uint4 e4, auto pf, uint4& slot) {
                             force_group_sgn<true, LT>(
                                 e4, pf, slot, rec, atype, off, olo, ui,
                                 0.5f * dFi, S, FR, it, it * nt, plen, ta,
                                 fx, fy, fz, ep, vir, vlo, r2min);
                           });
          else
            walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                           [&](uint4 e4, auto pf, uint4& slot) {
                             force_group_sgn<false, LT>(
                                 e4, pf, slot, rec, atype, off, olo, ui,
                                 0.5f * dFi, S, FR, it, it * nt, plen, ta,
                                 fx, fy, fz, ep, vir, vlo, r2min);
                           });
          is_short |= (r2min < pb);
        } else if constexpr (MODE == FM_SG && !CUB && !EE) {
          // one grid: sums in units of 1/(2 istep), scaled once per slot
          R r2min = (R)3.0e38;
          if constexpr (sizeof(R) == 4) {
            const float4* __restrict__ Q =
                reinterpret_cast<const float4*>(K + (a.joint - klo));
            const float2* __restrict__ Rr =
                reinterpret_cast<const float2*>(K + (a.joint - klo) + 2 * plen);
            const TArg ta = targ(tp);
            walk_either<L, NTL>(w16, a.nbr, i, a.kmax, gs, a.nnb[i],
                           [&](uint4 e4, auto pf, uint4& slot) {
                             force_group_sgf(e4, pf, slot, rec, off, olo, ui,
                                             0.5f * dFi, Q, Rr, ta, fx, fy,
                                             fz, ep, vir, vlo);
                           });
          } else {
            const R4* __restrict__ S =
                reinterpret_cast<const R4*>(K + (a.joint - klo));
            const R* __restrict__ Sr =
                reinterpret_cast<const R*>(K + (a.joint - klo) + 2 * plen);
            list_walk<L>(nw, a.nnb[i],
                         [&](uint4 e4, ListRef<L> pf, uint4& slot) {
                           force_group_sg<R>(e4, pf, slot, rec, off, ui,
                                             (R)0.5 * dFi, S, Sr, plen, pb,
                                             pe, pis, fx, fy, fz, ep, vir,
                                             r2min);
                         });
          }
          is_short |= (r2min < pb);
        } else {
          ForceCols<R> fc = {pk0, plen, pb, pe, pis, p2is,
                             rk0, rlen, rb, re, ris, r2is};
          const R dmi = EE ? a.eedm[i] : (R)0;
          list_walk<L>(nw, a.nnb[i],
                       [&](uint4 e4, ListRef<L> pf, uint4& slot) {
                         force_group<R, NT1, HP, HE, CUB, EE>(
                             e4, pf, slot, rec, atype, off, ui, dFi, it, nt,
                             K, cpp, cpr, fc, a.eedm, dmi, fx, fy, fz, ep,
                             vir, is_short);
                       });
        }
      }
      fx = lane_group_sum<L>(fx);
      fy = lane_group_sum<L>(fy);
      fz = lane_group_sum<L>(fz);
      ep = lane_group_sum<L>(ep);
      vir = lane_group_sum<L>(vir);
      if (sizeof(R) == 4) vlo = lane_group_sum<L>(vlo);
      if (gs == p) {
        kx = fx;
        ky = fy;
        kz = fz;
        kep = ep;
        kvir = vir;
        kvlo = vlo;
      }
    }
    // the one-grid loops sum in units of 1/(2 istep); the virial is scaled
    // in FP64 with its remainder sum added
    constexpr bool ONEGRID =
        (MODE == FM_SG || (MODE == FM_SGN && sizeof(R) == 4)) && !CUB && !EE;
    if (ONEGRID) {
      kx *= p2is;
      ky *= p2is;
      kz *= p2is;
    }
    kx = to_slot_lane<L>(kx, lane);
    ky = to_slot_lane<L>(ky, lane);
    kz = to_slot_lane<L>(kz, lane);
    kep = to_slot_lane<L>(kep, lane);
    kvir = to_slot_lane<L>(kvir, lane);
    if (sizeof(R) == 4) kvlo = to_slot_lane<L>(kvlo, lane);
    int i = c * IMDG_CHUNK + lane;
    double v[PB_N] = {0, 0, 0, 0, 0, 0};
    if (i - Rp.atom_off < Rp.natoms) {
      if (a.store) {
        // per-atom energy: half of every pair term plus F(rho_i)
        R ei = (R)0.5 * kep + (HE ? a.epot[i] : (R)0);
        a.epot[i] = ei;
        R4 fv;
        fv.x = kx;
        fv.y = ky;
        fv.z = kz;
        fv.w = (R)0;
        a.frc[i] = fv;
      }
      v[PB_EPAIR] = 0.5 * (double)kep;
      v[PB_VIRIAL] = ONEGRID ? -0.5 * ((double)kvir + (double)kvlo) *
                                   (double)p2is
                             : -0.5 * (double)kvir;
      if (INTEGRATE)
        integrate_slot(a, Rp, r, i, (double)kx, (double)ky, (double)kz, v);
    }
    put_part(a.partB, a.part_chunk, c, v, acc);
  }
  flush_part(a.partB, a.part_chunk, acc);
  flag_short(a.ctl, is_short);
  if (a.final_inline) last_block_finalize(a, true, INTEGRATE);
}

// move_atoms + check_nblist on the stored forces (the unfused entry point).
template <typename R>
__global__ void __launch_bounds__(IMDG_STEP_THREADS) k_move(StepArgs<R> a) {
  typedef typename V4<R>::T R4;
  if (a.ctl->frozen) return;
  Sched S = make_sched(a.nchunk, a.xcd_remap);
  const int lane = threadIdx.x & (IMDG_WAVE - 1);
  double acc[PB_N] = {0, 0, 0, 0, 0, 0};
  for (int c = S.c; c < S.c1; c += S.stride) {
    int r = a.blk_rep[c / (IMDG_BLOCK / IMDG_CHUNK)];
    const RepDev& Rp = a.rep[r];
    int i = c * IMDG_CHUNK + lane;
    double v[PB_N] = {0, 0, 0, 0, 0, 0};
    if (i - Rp.atom_off < Rp.natoms) {
      R4 f = a.frc[i];
      integrate_slot(a, Rp, r, i, (double)f.x, (double)f.y, (double)f.z, v);
    }
    put_part(a.partB, a.part_chunk, c, v, acc);
  }
  flush_part(a.partB, a.part_chunk, acc);
  if (a.final_inline) last_block_finalize(a, false, true);
}

// The end of a step as a launch of its own (one block), when the step
// kernels leave it out (final_inline = 0).
template <typename R>
__global__ void __launch_bounds__(IMDG_STEP_THREADS)
    k_finalize(StepArgs<R> a, int forces, int moved) {
  if (a.ctl->frozen) return;
  finalize_step(a, forces != 0, moved != 0, blockIdx.x, gridDim.x);
}

// Pair-only systems: the force loop gathers rec_b, which holds a copy of the
// (u, type) records (rec_a is rewritten by the integration in that loop).
template <typename R>
__global__ void __launch_bounds__(256) k_copyrec(StepArgs<R> a, int nslots) {
  if (a.ctl->frozen) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) a.rec_b[i] = a.rec_a[i];
}

// Per-atom stress tensor of the step just done (STRESS_TENS:
// src/imd_forces_nbl.c:558-581, 1283-1318, src/imd_integrate.c:411-431,
// 1085-1094): S_i = -1/2 sum_j g_ij d d^T over the full list, with d = x_j -
// x_i and g_ij the pair gradient of the force loop (pair term plus
// 1/2 (2F'_i rho'_j + 2F'_j rho'_i)), plus p_i p_i^T / m_i with the
// momenta after the step; components xx yy zz yz zx xy.  It gathers rec_b,
// which still holds the (u, 2F') records the step's force loop read (the
// integration rewrites rec_a only), on the generic knot tables.  Launched
// only on steps whose stress is read.  One thread per slot in
// IMDG_BLOCK-thread blocks (one replica each); each block also writes the
// sums of its atoms' six components to spart (FP64).
template <typename R>
__global__ void __launch_bounds__(IMDG_BLOCK)
    k_stress(StepArgs<R> a, const TabDev<R>* __restrict__ tabs,
             double* __restrict__ stress, double* __restrict__ spart) {
  typedef typename V4<R>::T R4;
  typedef typename V2<R>::T R2;
  __shared__ R4 off[IMDG_NCODES];
  __shared__ double red[IMDG_NWAVE][6];
  __shared__ int sbase[IMDG_NWAVE][IMDG_META];
  if (a.ctl->frozen) return;
  const int b = blockIdx.x;
  const RepDev& Rp = a.rep[a.blk_rep[b]];
  if (threadIdx.x < IMDG_NCODES) {
    R4 o;
    o.x = Rp.offs[threadIdx.x][0];
    o.y = Rp.offs[threadIdx.x][1];
    o.z = Rp.offs[threadIdx.x][2];
    o.w = 0;
    off[threadIdx.x] = o;
  }
  // the 16-bit list format: base slots and format of the block's chunks
  const int wv = threadIdx.x / IMDG_WAVE, lane = threadIdx.x & (IMDG_WAVE - 1);
  if (a.list16 && lane <= IMDG_META_WIDE)
    sbase[wv][lane] =
        a.cmeta[(size_t)(b * IMDG_NWAVE + wv) * IMDG_META + lane];
  __syncthreads();
  const bool w16 = a.list16 && sbase[wv][IMDG_META_WIDE] == 0;
  const TabDev<R>& tp = tabs[0];
  const TabDev<R>& tr = tabs[1];
  const bool hp = a.has_pair != 0, he = a.has_eam != 0;
  const int nt = a.ntypes;
  const R2* __restrict__ K = a.knots;
  const R4* __restrict__ rec = a.rec_b;
  const bool cub = a.cubic != 0;
  // twice the derivative of a table at segment knot k0 (3-point knots or
  // cubic segments)
  auto grad_at = [&](int k0, R chi, R twoistep) {
    return cub ? cub_grad<R>(K[k0], K[k0 + 1], chi, twoistep)
               : knot_grad<R>(K[k0], K[k0 + 1], chi, twoistep);
  };
  auto val_at = [&](int k0, R chi) {
    return cub ? cub_val<R>(K[k0], K[k0 + 1], chi)
               : knot_val<R>(K[k0], K[k0 + 1], chi);
  };
  const bool ee = a.eeam != 0;
  const int kmul = cub ? 2 : 1;
  const int i = b * IMDG_BLOCK + threadIdx.x;
  double s[6] = {0, 0, 0, 0, 0, 0};
  if (i - Rp.atom_off < Rp.natoms) {
    const R4 ui = rec[i];
    const int it = a.atype[i];
    const R dFi = he ? ui.w : (R)0;
    const R dmi = ee ? a.eedm[i] : (R)0;
    R t[6] = {0, 0, 0, 0, 0, 0};
    const uint4* lw = list_of(a.nbr, i, a.kmax);
    const char* lw16 = list16_of(a.nbr, i, a.kmax);
    const int ngroup = (a.nnb[i] + IMDG_GROUP - 1) / IMDG_GROUP;
    for (int gq = 0; gq < ngroup; gq++) {
      uint4 e4;
      if (w16)
        e4 = unpack16<1>(make_uint4(
            reinterpret_cast<const uint2*>(lw16 + (size_t)gq * 512)->x,
            reinterpret_cast<const uint2*>(lw16 + (size_t)gq * 512)->y, 0u,
            0u));
      else
        e4 = lw[(size_t)gq * IMDG_WAVE];
      uint32_t en[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        // entries past the list's end are padding: a far-away image,
        // outside every table
        const uint32_t code =
            w16 ? en[u] >> IMDG_K16_BITS : en[u] >> IMDG_IDX_BITS;
        const int j = w16 ? sbase[wv][code] + (int)(en[u] & IMDG_K16_MASK)
                          : (int)(en[u] & IMDG_IDX_MASK);
        const R4 uj = rec[j];
        const R4 o = off[code];
        const R dx = (uj.x - ui.x) + o.x;
        const R dy = (uj.y - ui.y) + o.y;
        const R dz = (uj.z - ui.z) + o.z;
        const R r2 = dx * dx + dy * dy + dz * dz;
        const int jt = a.atype[j];
        const int col = it * nt + jt, colT = jt * nt + it;
        R g = (R)0;
        if (hp && r2 <= tp.end[col]) {
          R chi;
          int k = seg_locate(r2, tp.begin[col], tp.end[col], tp.invstep[col],
                             tp.len[col], chi);
          g += grad_at(tp.off[col] + kmul * k, chi, tp.twoistep[col]);
        }
        if (he && (r2 < tr.end[colT] || r2 < tr.end[col])) {
          // rho'_i from column jt*nt+it, rho'_j from it*nt+jt
          R chi;
          int k = seg_locate(r2, tr.begin[colT], tr.end[colT],
                             tr.invstep[colT], tr.len[colT], chi);
          R rpi = grad_at(tr.off[colT] + kmul * k, chi, tr.twoistep[colT]);
          R rvi = ee ? val_at(tr.off[colT] + kmul * k, chi) : (R)0;
          R rpj = rpi, rvj = rvi;
          if (colT != col) {
            k = seg_locate(r2, tr.begin[col], tr.end[col], tr.invstep[col],
                           tr.len[col], chi);
            rpj = grad_at(tr.off[col] + kmul * k, chi, tr.twoistep[col]);
            if (ee) rvj = val_at(tr.off[col] + kmul * k, chi);
          }
          g += (R)0.5 * (dFi * rpj + uj.w * rpi);
          if (ee) g += dmi * rvj * rpj + a.eedm[j] * rvi * rpi;
        }
        t[0] = fma(dx * dx, g, t[0]);
        t[1] = fma(dy * dy, g, t[1]);
        t[2] = fma(dz * dz, g, t[2]);
        t[3] = fma(dy * dz, g, t[3]);
        t[4] = fma(dz * dx, g, t[4]);
        t[5] = fma(dx * dy, g, t[5]);
      }
    }
    const double4 p = a.mom[a.ctl->pos_cur][i];
    const double m = p.w;
    s[0] = -0.5 * (double)t[0] + p.x * p.x / m;
    s[1] = -0.5 * (double)t[1] + p.y * p.y / m;
    s[2] = -0.5 * (double)t[2] + p.z * p.z / m;
    s[3] = -0.5 * (double)t[3] + p.y * p.z / m;
    s[4] = -0.5 * (double)t[4] + p.z * p.x / m;
    s[5] = -0.5 * (double)t[5] + p.x * p.y / m;
#pragma unroll
    for (int q = 0; q < 6; q++) stress[(size_t)i * 6 + q] = s[q];
  }
  const int w = threadIdx.x / IMDG_WAVE;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    double v = wave_sum(s[q]);
    if ((threadIdx.x & (IMDG_WAVE - 1)) == 0) red[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = 0.0;
    for (int k = 0; k < IMDG_NWAVE; k++) v += red[k][threadIdx.x];
    spart[(size_t)b * 6 + threadIdx.x] = v;
  }
}

// ---------------------------------------------------------------------------
// launch plans and launchers
// ---------------------------------------------------------------------------

template <typename R, bool NT1, bool LT, bool CUB, bool EE, bool NTL>
const void* density_fn_l(int lpa) {
  if (lpa == 4)
    return reinterpret_cast<const void*>(
        &k_density<R, NT1, LT, 4, CUB, EE, NTL>);
  if (lpa == 2)
    return reinterpret_cast<const void*>(
        &k_density<R, NT1, LT, 2, CUB, EE, NTL>);
  return reinterpret_cast<const void*>(&k_density<R, NT1, LT, 1, CUB, EE, NTL>);
}

// ntl: the list is loaded non-temporally (StepArgs::list_nt); only the
// variants with FP32 one-grid walks have such an instance, the others load
// the list words with the default policy whatever ntl says.
template <typename R, bool NT1, bool LT, bool CUB, bool EE>
const void* density_fn(int lpa, bool ntl) {
  constexpr bool ONEGRID = sizeof(R) == 4 && !CUB && !EE;
  if constexpr (ONEGRID)
    if (ntl) return density_fn_l<R, NT1, LT, CUB, EE, true>(lpa);
  return density_fn_l<R, NT1, LT, CUB, EE, false>(lpa);
}

template <typename R, int MODE, bool INTEGRATE, bool HP, bool HE, bool LT,
          bool CUB, bool EE, bool NTL>
const void* force_fn4(int lpa) {
  if (lpa == 4)
    return reinterpret_cast<const void*>(
        &k_force<R, MODE, INTEGRATE, HP, HE, LT, 4, CUB, EE, NTL>);
  if (lpa == 2)
    return reinterpret_cast<const void*>(
        &k_force<R, MODE, INTEGRATE, HP, HE, LT, 2, CUB, EE, NTL>);
  return reinterpret_cast<const void*>(
      &k_force<R, MODE, INTEGRATE, HP, HE, LT, 1, CUB, EE, NTL>);
}

// ntl as for density_fn; the non-temporal instances exist for the step
// (INTEGRATE) only.
template <typename R, int MODE, bool INTEGRATE, bool HP, bool HE, bool LT,
          bool CUB, bool EE>
const void* force_fn3(int lpa, bool ntl) {
  constexpr bool ONEGRID = sizeof(R) == 4 && INTEGRATE && !CUB && !EE &&
                           (MODE == FM_SG || MODE == FM_SGN);
  if constexpr (ONEGRID)
    if (ntl)
      return force_fn4<R, MODE, INTEGRATE, HP, HE, LT, CUB, EE, true>(lpa);
  return force_fn4<R, MODE, INTEGRATE, HP, HE, LT, CUB, EE, false>(lpa);
}

template <typename R, int MODE, bool INTEGRATE, bool LT, bool CUB, bool EE>
const void* force_fn2(bool hp, bool he, int lpa, bool ntl) {
  if constexpr (MODE == FM_SG || MODE == FM_SGN) {
    return force_fn3<R, MODE, INTEGRATE, true, true, LT, CUB, EE>(lpa, ntl);
  } else {
    if (hp && he)
      return force_fn3<R, MODE, INTEGRATE, true, true, LT, CUB, EE>(lpa, ntl);
    if (EE) return nullptr;  // EEAM needs the EAM tables
    if (hp)
      return force_fn3<R, MODE, INTEGRATE, true, false, LT, CUB, EE>(lpa,
                                                                     ntl);
    return force_fn3<R, MODE, INTEGRATE, false, true, LT, CUB, EE>(lpa, ntl);
  }
}

template <typename R, int MODE>
const void* force_fn1(bool integrate, bool lt, bool hp, bool he, int lpa,
                      bool ntl) {
  if (integrate)
    return lt ? force_fn2<R, MODE, true, true, false, false>(hp, he, lpa, ntl)
              : force_fn2<R, MODE, true, false, false, false>(hp, he, lpa,
                                                              ntl);
  return lt ? force_fn2<R, MODE, false, true, false, false>(hp, he, lpa, ntl)
            : force_fn2<R, MODE, false, false, false, false>(hp, he, lpa,
                                                             ntl);
}

// The table variants (cubic segments of the 4-point and spline
// interpolations, EEAM) take the generic loops with the tables in HBM/L2.
template <typename R, int MODE, bool CUB, bool EE>
const void* force_fn1_var(bool integrate, bool hp, bool he, int lpa,
                          bool ntl) {
  return integrate
             ? force_fn2<R, MODE, true, false, CUB, EE>(hp, he, lpa, ntl)
             : force_fn2<R, MODE, false, false, CUB, EE>(hp, he, lpa, ntl);
}

template <typename R, int MODE>
const void* force_fn_var(bool integrate, bool hp, bool he, int lpa, bool ntl,
                         bool cub, bool ee) {
  if (cub && ee)
    return force_fn1_var<R, MODE, true, true>(integrate, hp, he, lpa, ntl);
  if (cub)
    return force_fn1_var<R, MODE, true, false>(integrate, hp, he, lpa, ntl);
  return force_fn1_var<R, MODE, false, true>(integrate, hp, he, lpa, ntl);
}

// mode: 0 one species, 1 one species on a shared grid (needs hp and he),
// 2 several species; lpa: lanes per slot (1, 2 or 4); cub: cubic-segment
// tables, ee: EEAM (both: modes 0 and 2 only, lt false)
template <typename R>
const void* force_fn(int mode, bool integrate, bool lt, bool hp, bool he,
                     int lpa, bool ntl, bool cub, bool ee) {
  if (cub || ee) {
    if (lt || mode == FM_SG) return nullptr;
    if (mode == FM_NT1)
      return force_fn_var<R, FM_NT1>(integrate, hp, he, lpa, ntl, cub, ee);
    return force_fn_var<R, FM_NTN>(integrate, hp, he, lpa, ntl, cub, ee);
  }
  if (mode == FM_SG)
    return force_fn1<R, FM_SG>(integrate, lt, true, true, lpa, ntl);
  if (mode == FM_SGN) {
    // FP32 only; lt: the per-type rho segments in LDS
    if constexpr (sizeof(R) != 4) {
      return nullptr;
    } else {
      if (lt)
        return integrate
                   ? force_fn2<R, FM_SGN, true, true, false, false>(
                         true, true, lpa, ntl)
                   : force_fn2<R, FM_SGN, false, true, false, false>(
                         true, true, lpa, ntl);
      return integrate
                 ? force_fn2<R, FM_SGN, true, false, false, false>(true, true,
                                                                   lpa, ntl)
                 : force_fn2<R, FM_SGN, false, false, false, false>(true, true,
                                                                    lpa, ntl);
    }
  }
  if (mode == FM_NT1)
    return force_fn1<R, FM_NT1>(integrate, lt, hp, he, lpa, ntl);
  return force_fn1<R, FM_NTN>(integrate, lt, hp, he, lpa, ntl);
}

template <typename R, bool CUB, bool EE>
const void* density_fn_var(bool nt1, int lpa, bool ntl) {
  return nt1 ? density_fn<R, true, false, CUB, EE>(lpa, ntl)
             : density_fn<R, false, false, CUB, EE>(lpa, ntl);
}

template <typename R>
const void* density_fn(bool nt1, bool lt, int lpa, bool ntl, bool cub,
                       bool ee) {
  if (cub || ee) {
    if (lt) return nullptr;
    if (cub && ee) return density_fn_var<R, true, true>(nt1, lpa, ntl);
    if (cub) return density_fn_var<R, true, false>(nt1, lpa, ntl);
    return density_fn_var<R, false, true>(nt1, lpa, ntl);
  }
  if (nt1)
    return lt ? density_fn<R, true, true, false, false>(lpa, ntl)
              : density_fn<R, true, false, false, false>(lpa, ntl);
  return lt ? density_fn<R, false, true, false, false>(lpa, ntl)
            : density_fn<R, false, false, false, false>(lpa, ntl);
}

// Persistent grid of a kernel: resident blocks per CU (from the occupancy
// calculator with the launch's dynamic LDS) times the CU count, rounded
// down to a multiple of the 8 XCDs; 0 if the kernel cannot be launched with
// that much dynamic LDS.
static int persistent_grid(const void* fn, int threads, size_t lds, int ncu) {
  if (!fn) return 0;  // no such variant
  hipFuncAttributes fa;
  if (hipFuncGetAttributes(&fa, fn) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  int dev = 0, maxlds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (fa.sharedSizeBytes + lds > (size_t)maxlds) return 0;
  if (lds > 0 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) !=
      hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (per < 1) return 0;
  int g = per * ncu;
  if (g >= 8) g &= ~7;
  return g;
}

// The grid of a step loop fits both list policies' instances (and sets
// the dynamic-LDS limit of both): list_nt may change with every rebuild.
template <typename R>
int step_grid_density(bool nt1, bool lt, size_t lds, int ncu, int lpa,
                      int threads, bool cub, bool ee) {
  int g[2];
  for (int q = 0; q < 2; q++)
    g[q] = persistent_grid(density_fn<R>(nt1, lt, lpa, q != 0, cub, ee),
                           threads, lds, ncu);
  return g[0] < g[1] ? g[0] : g[1];
}

template <typename R>
int force_threads(int mode) {
  return mode == FM_SG    ? ForceThreads<R, FM_SG>::value
         : mode == FM_SGN ? ForceThreads<R, FM_SGN>::value
         : mode == FM_NT1 ? ForceThreads<R, FM_NT1>::value
                          : ForceThreads<R, FM_NTN>::value;
}

template <typename R>
int step_grid_force(int mode, bool integrate, bool lt, bool hp, bool he,
                    size_t lds, int ncu, int lpa, int threads, bool cub,
                    bool ee) {
  if (threads > force_threads<R>(mode)) return 0;
  int g[2];
  for (int q = 0; q < 2; q++)
    g[q] = persistent_grid(
        force_fn<R>(mode, integrate, lt, hp, he, lpa, q != 0, cub, ee),
        threads, lds, ncu);
  return g[0] < g[1] ? g[0] : g[1];
}

template <typename R>
int step_grid_move(int ncu) {
  return persistent_grid(reinterpret_cast<const void*>(&k_move<R>),
                         IMDG_STEP_THREADS, 0, ncu);
}

template <typename R>
hipError_t launch_density(const StepArgs<R>& a, const TabDev<R>* tabs,
                          bool nt1, int grid, hipStream_t st) {
  bool lt = a.lds_n > 0;
  size_t lds = (size_t)a.lds_n * sizeof(typename V2<R>::T);
  struct {
    StepArgs<R> a;
    const TabDev<R>* t;
  } pk = {a, tabs};
  void* kargs[] = {&pk.a, &pk.t};
  const void* fn =
      density_fn<R>(nt1, lt, a.lpa_den, a.list_nt != 0, a.cubic != 0,
                    a.eeam != 0);
  if (!fn) return hipErrorInvalidDeviceFunction;
  return hipLaunchKernel(fn, dim3(grid), dim3(a.thr_den), kargs, lds, st);
}

template <typename R>
hipError_t launch_force(const StepArgs<R>& a, const TabDev<R>* tabs,
                        int mode, bool integrate, int grid, hipStream_t st) {
  bool lt = a.lds_n > 0;
  size_t lds = (size_t)a.lds_n * sizeof(typename V2<R>::T);
  StepArgs<R> aa = a;
  const TabDev<R>* tt = tabs;
  void* kargs[] = {&aa, &tt};
  const void* fn =
      force_fn<R>(mode, integrate, lt, a.has_pair != 0, a.has_eam != 0,
                  a.lpa_frc, a.list_nt != 0, a.cubic != 0, a.eeam != 0);
  if (!fn) return hipErrorInvalidDeviceFunction;
  return hipLaunchKernel(fn, dim3(grid), dim3(a.thr_frc), kargs, lds, st);
}

template <typename R>
hipError_t launch_move(const StepArgs<R>& a, int grid, hipStream_t st) {
  StepArgs<R> aa = a;
  void* kargs[] = {&aa};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_move<R>),
                         dim3(grid), dim3(IMDG_STEP_THREADS), kargs, 0, st);
}

template <typename R>
hipError_t launch_finalize(const StepArgs<R>& a, bool forces, bool moved,
                           hipStream_t st) {
  // several replicas: one block per IMDG_STEP_WAVES replicas (one wavefront
  // each); one replica: one wavefront (its lanes sum the block partials, and
  // no combine over wavefronts follows: the end of the step is a chain of
  // dependent operations, and a shorter chain ends it sooner)
  int nblk =
      a.part_chunk ? (a.nrep + IMDG_STEP_WAVES - 1) / IMDG_STEP_WAVES : 1;
  int threads = a.part_chunk ? IMDG_STEP_THREADS : IMDG_WAVE;
  hipLaunchKernelGGL((k_finalize<R>), dim3(nblk), dim3(threads), 0, st, a,
                     forces ? 1 : 0, moved ? 1 : 0);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_copyrec(const StepArgs<R>& a, int nslots, hipStream_t st) {
  hipLaunchKernelGGL((k_copyrec<R>), dim3((nslots + 255) / 256), dim3(256), 0,
                     st, a, nslots);
  return hipGetLastError();
}

// nblk: IMDG_BLOCK-slot blocks of all replicas; stress holds 6 doubles per
// slot, spart 6 per block.
template <typename R>
hipError_t launch_stress(const StepArgs<R>& a, const TabDev<R>* tabs,
                         double* stress, double* spart, int nblk,
                         hipStream_t st) {
  hipLaunchKernelGGL((k_stress<R>), dim3(nblk), dim3(IMDG_BLOCK), 0, st, a,
                     tabs, stress, spart);
  return hipGetLastError();
}

#define IMDG_INSTANTIATE(R)                                                   \
  template int step_grid_density<R>(bool, bool, size_t, int, int, int,     \
                                    bool, bool);                              \
  template int step_grid_force<R>(int, bool, bool, bool, bool, size_t, int,  \
                                  int, int, bool, bool);                      \
  template int force_threads<R>(int);                                         \
  template int step_grid_move<R>(int);                                        \
  template hipError_t launch_density<R>(const StepArgs<R>&,                   \
                                        const TabDev<R>*, bool, int,          \
                                        hipStream_t);                         \
  template hipError_t launch_force<R>(const StepArgs<R>&, const TabDev<R>*,   \
                                      int, bool, int, hipStream_t);           \
  template hipError_t launch_move<R>(const StepArgs<R>&, int, hipStream_t);   \
  template hipError_t launch_copyrec<R>(const StepArgs<R>&, int, hipStream_t); \
  template hipError_t launch_stress<R>(const StepArgs<R>&, const TabDev<R>*,  \
                                       double*, double*, int, hipStream_t);   \
  template hipError_t launch_finalize<R>(const StepArgs<R>&, bool, bool,     \
                                         hipStream_t);

// One precision per object when IMDG_STEP_FP32 or IMDG_STEP_FP64 is
// defined (the build compiles this file once per type, in parallel).
#if defined(IMDG_STEP_FP32) || !defined(IMDG_STEP_FP64)
IMDG_INSTANTIATE(float)
#endif
#if defined(IMDG_STEP_FP64) || !defined(IMDG_STEP_FP32)
IMDG_INSTANTIATE(double)
#endif

}  // namespace imdg
