// imdg_noalloc.hip -- check_nblist on caller-owned device arrays, with no
// allocation: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight loops of a host program that owns its positions on the
// device).
//
// imdg_check_nblist (the context's skin test, src/imd_forces_nbl.c:
// 2007-2037) reduces per-block maxima of |x - x_nbl|^2 in a buffer the
// context allocates.  This variant takes the positions and the positions
// at the last list build as device pointers, reduces through a file-scope
// device array instead (`scratch`: one double per block of IMDG_NA_BLOCK
// atoms, IMDG_NA_MAXBLK blocks), and leaves the maximum and IMD's decision
// max > (nbl_margin / 2)^2 in device memory the caller passes.  Two kernel
// launches on the caller's stream, no host synchronisation, no hipMalloc,
// hipFree or allocating library call.
//
// Every call on one GPU uses that GPU's one `scratch` array: calls that
// overlap on one GPU (two streams, two host threads) race.  Each call
// writes every scratch entry it later reads, so nothing needs clearing.
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdio.h>

#include "../../include/imdg.h"
#include "imdg_device.h"

#define IMDG_NA_BLOCK 256
#define IMDG_NA_MAXBLK 65536
#define IMDG_NA_MAXDEV 64

namespace imdg {

// Per-block maxima of the displacement since the last list build.
static __device__ __align__(256) double scratch[IMDG_NA_MAXBLK];

__global__ void __launch_bounds__(IMDG_NA_BLOCK)
    k_disp_partial(const double* __restrict__ x,
                   const double* __restrict__ xr, long n, double* part) {
#pragma clang fp contract(off)
  __shared__ double red[IMDG_NA_BLOCK / IMDG_WAVE];
  const long i = (long)blockIdx.x * IMDG_NA_BLOCK + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    // check_nblist's squared displacement, operations rounded one by one
    double d0 = x[3 * i] - xr[3 * i];
    double d1 = x[3 * i + 1] - xr[3 * i + 1];
    double d2 = x[3 * i + 2] - xr[3 * i + 2];
    v = d0 * d0 + d1 * d1 + d2 * d2;
  }
  v = wave_max(v);
  if ((threadIdx.x & (IMDG_WAVE - 1)) == 0) red[threadIdx.x / IMDG_WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < IMDG_NA_BLOCK / IMDG_WAVE; k++) m = fmax(m, red[k]);
    part[blockIdx.x] = m;
  }
}

__global__ void __launch_bounds__(1024)
    k_disp_final(const double* __restrict__ part, int nb, double h2,
                 double* max_disp2, int* invalid) {
  __shared__ double red[1024 / IMDG_WAVE];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, part[b]);
  m = wave_max(m);
  if ((threadIdx.x & (IMDG_WAVE - 1)) == 0) red[threadIdx.x / IMDG_WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x / IMDG_WAVE); k++) m = fmax(m, red[k]);
    m = fmax(m, red[0]);
    if (max_disp2) *max_disp2 = m;
    if (invalid) *invalid = m > h2 ? 1 : 0;
  }
}

// Device address of `scratch` on each GPU (each GPU holds its own copy),
// looked up once per device with that device current.
static double* g_scratch[IMDG_NA_MAXDEV];
static std::once_flag g_once[IMDG_NA_MAXDEV];
static hipError_t g_lookup[IMDG_NA_MAXDEV];
static thread_local char g_msg[256];

static int na_fail(int rc, const char* msg) {
  snprintf(g_msg, sizeof g_msg, "%s", msg);
  return rc;
}

// The message of the last failed call on this host thread ("" if none):
// imdg_last_error(NULL) returns it.
const char* noalloc_message() { return g_msg; }

}  // namespace imdg

using namespace imdg;

extern "C" int imdg_check_nblist_noalloc(int device, const double* x,
                                         const double* x_nbl, long n,
                                         double nbl_margin, double* max_disp2,
                                         int* invalid, void* stream) {
  g_msg[0] = 0;
  if (device < 0 || device >= IMDG_NA_MAXDEV)
    return na_fail(IMDG_ERR_ARG, "imdg_check_nblist_noalloc: device out of "
                                 "range");
  if (n < 0 || (n > 0 && (!x || !x_nbl)) || !(nbl_margin >= 0) ||
      (!max_disp2 && !invalid))
    return na_fail(IMDG_ERR_ARG, "imdg_check_nblist_noalloc: invalid "
                                 "argument");
  if (n > (long)IMDG_NA_MAXBLK * IMDG_NA_BLOCK) {
    snprintf(g_msg, sizeof g_msg,
             "imdg_check_nblist_noalloc: %ld atoms exceed its fixed scratch "
             "(%d blocks of %d atoms: at most %ld atoms)", n, IMDG_NA_MAXBLK,
             IMDG_NA_BLOCK, (long)IMDG_NA_MAXBLK * IMDG_NA_BLOCK);
    return IMDG_ERR_ARG;
  }
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess)
    return na_fail(IMDG_ERR_HIP, "imdg_check_nblist_noalloc: hipGetDevice "
                                 "failed");
  if (prev != device && hipSetDevice(device) != hipSuccess)
    return na_fail(IMDG_ERR_HIP, "imdg_check_nblist_noalloc: hipSetDevice "
                                 "failed");
  std::call_once(g_once[device], [device] {
    void* p = nullptr;
    g_lookup[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch[device] = (double*)p;
  });
  int rc = IMDG_OK;
  if (g_lookup[device] != hipSuccess || !g_scratch[device]) {
    rc = na_fail(IMDG_ERR_HIP, "imdg_check_nblist_noalloc: "
                               "hipGetSymbolAddress(scratch) failed");
  } else {
    hipStream_t st = (hipStream_t)stream;
    const int nb = (int)((n + IMDG_NA_BLOCK - 1) / IMDG_NA_BLOCK);
    const double h = 0.5 * nbl_margin;
    if (nb > 0)
      hipLaunchKernelGGL(k_disp_partial, dim3(nb), dim3(IMDG_NA_BLOCK), 0, st,
                         x, x_nbl, n, g_scratch[device]);
    hipLaunchKernelGGL(k_disp_final, dim3(1), dim3(1024), 0, st,
                       (const double*)g_scratch[device], nb, h * h, max_disp2,
                       invalid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      snprintf(g_msg, sizeof g_msg, "imdg_check_nblist_noalloc: launch "
               "failed: %s", hipGetErrorString(e));
      rc = IMDG_ERR_HIP;
    }
  }
  if (prev != device) hipSetDevice(prev);
  return rc;
}
