// C-ABI layer of libwql3d_hip.so: runtime, block lifecycle (device mirrors),
// upload/download, and the stage orchestration of time_step_RK
// (time_step.f90:68-150).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <thread>
#include <vector>

#include "wql3d_adjoint.h"
#include "wql3d_internal.h"

using namespace wql;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(WQL_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                   \
  } while (0)

// Kennedy-Carpenter (5,4) 2N-storage coefficients (time_step.f90:53-62)
const double kRkA[5] = {0.0, -567301805773.0 / 1357537059087.0,
                        -2404267990393.0 / 2016746695238.0, -3550918686646.0 / 2091501179385.0,
                        -1275806237668.0 / 842570457699.0};
const double kRkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double kRkC[6] = {0.0,
                        1432997174477.0 / 9575080441755.0,
                        2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0,
                        2802321613138.0 / 2924317926251.0,
                        1.0};

long long box_extent(const wql_box& b, int d) { return (long long)b.hi[d] - b.lo[d] + 1; }

}  // namespace

struct wql_block {
  wql_block_desc d;
  int device;                // HIP device current at wql_block_create
  int nx, ny, nz;
  // plane stride: nodes rounded up to a multiple of kPlaneAlign, so every
  // plane starts on a 1 KiB (fp32) / 2 KiB (fp64) boundary and the streaming
  // kernels' wavefront accesses never straddle extra cache lines
  long long P;
  size_t esz;                // sizeof(T) of geometry/material/memory arrays
  size_t ssz;                // sizeof(S) of the RK state F, DF (8, or 4 in FP32-Kahan)
  int cur;                   // current F buffer
  void* F[2];                // S planes
  void* DF;
  void* Fc;                  // FP32-Kahan: compensation of F (float planes)
  void* Fo[2];               // FP32-mixed: F[k] rounded to float (stencil operands)
  bool fo_valid;             // Fo[cur] holds F[cur] rounded (cleared wherever F changes
                             // outside the fused stage)
  void* G;
  void* M;
  void* eta;
  void* Deta;
  void* Qs;
  void* Qp;
  void* pQ[6];
  void* pDQ[6];
  long long Ps[6];           // PML slab plane sizes
  int slab_lo[6][3];         // slab node origin (0-based)
  int slab_n[6][3];          // slab extents
  void* face[6];
  double* face64[6];         // host copy of face[f] in FP64 (the interface reads it)
  long long NF[6];
  hipStream_t stream;
  int nsrc;
  wql_source src[kMaxSources];
  int nrec;
  int nslots;
  long long* rec_idx;
  double* seis;
  double* X;      // node coordinates, 3 planes (wql_block_set_coordinates)
  int slot;
  long long bytes;
  int tune_nt, tune_xcd, tune_band, tune_split;  // node-kernel launch shape
  int tune_zm, tune_zchunk;                      // z-marching kernel, z levels per chunk
  int tune_mem_npt;                              // nodes per thread of k_memory_update
  int tune_mem_nt;                               // non-temporal k_memory_update streams
  int tune_mem_xcd;                              // its workgroup order over the XCDs
  int tune_adj_xcd;                              // adjoint node kernels' order
  void* arena;    // one allocation holding F[2], DF, G, M, eta, Deta, Qs, Qp, W
  void* W;        // 12 planes of T: split anelastic stage hand-off
  int mech_zero;  // mechanisms whose eta and Deta are known to be all zero
  // kernel timing of fused stages (wql_block_set_timing): 4 events per stage
  // around the interior, shell and memory-variable kernels
  int timing;
  int nev_used;
  int nev_alloc;
  hipEvent_t* ev;
  // discrete adjoint (wql_adjoint_create): adjoint of (F, eta) and of
  // (DF, Deta), gradient planes (rho, lambda, mu), workspace, state slots
  int adj_nslots;    // full state slots (F and eta)
  int adj_nslots_f;  // slots after those that hold F only
  double* adj_Fb;
  void* adj_Eb;
  double* adj_DFb;
  void* adj_DEb;
  double* adj_g;
  void* adj_Z;
  double* slot_F;
  void* slot_E;
  int* slot_mz;
  // two-block interface (wql_iface_create): the interface the block is
  // coupled through and the SAT forcing planes of its coupled face
  wql_iface* iface;
  double* iface_sat[6];
  // z-chunk pipeline of anelastic stages (env WQL3D_STAGE_CHUNKS, default
  // kDefaultChunks = 1, i.e. off: at 385^3 with Q8 and PML 2/4/8 chunks ran
  // 3%/6%/14% slower, since the interior kernel's two waves per SIMD fill the
  // register file and leave no room for memory-kernel waves on its CUs, so the
  // two kernels do not overlap): the stencil kernels of chunk k run on `stream`, the
  // memory-variable kernel of chunk k on `stream2` behind ev_sten[k]; the
  // stencil of chunk k of the next stage waits for ev_mem[k + 1] (its z halo).
  // `pending` = memory kernels on stream2 that `stream` has not yet waited for.
  int nchunks;
  hipStream_t stream2;
  hipEvent_t ev_sten[kMaxChunks];
  hipEvent_t ev_mem[kMaxChunks];
  int pending;
  // pinned host staging of host-array transfers (allocated on the first
  // one, one plane of doubles each): a component is packed into one buffer
  // while the other's copy runs; stage_ev[k] marks buffer k free again
  void* stage_buf[2];
  size_t stage_bytes;
  hipEvent_t stage_ev[2];
  int stage_k;
};

// Two-block q-interface (wql_iface_create): block L's x-high face against
// block R's x-low face, ny x nz nodes, node index y + ny*z.
// make the block's stream wait for the pipelined memory-variable kernels
static int join(wql_block* b);

struct wql_iface {
  wql_block* L;
  wql_block* R;
  int ny, nz;
  long long NF;
  int coupling;
  // one allocation of kIfacePlanes planes: S 4, DS 4, trup, prestress of L 3
  // and of R 3, fs, fd, Dc, SAT forcing of L 9 and of R 9, the rate-and-state
  // state W, its rate DW, the parameter a, the nucleation amplitude and the
  // flash-weakening velocity V_w (rate-weakening), and FP64 copies of the
  // coupled faces' 13 frame planes and 3 material planes of each side
  double* buf;
  double *S, *DS, *trup, *T0L, *T0R, *fs, *fd, *Dc, *satL, *satR;
  double *W, *DW, *rs_a, *rs_nuc, *rs_vw, *frL, *frR, *matL, *matR;
  double alpha;     // linear friction: tau = alpha v
  double* frames;   // nslots * kFramePlanes planes of fault output
  int nslots, slot;
  hipEvent_t ev_iface;  // on L's stream after k_iface: block R's stage waits for it
  hipEvent_t ev_right;  // on R's stream after block R's stage: the next k_iface waits
};

// ---------------------------------------------------------------------------
// close_domain's shutdown check (domain.f90:520-545) on the device: the
// largest |F| and |eta| of a block and the count of non-finite values.  Two
// passes: each workgroup of k_max_state_part reduces a grid-stride share of
// every plane to one partial (3 doubles), then one workgroup of
// k_max_state_final reduces the partials.  Only the nodes of each plane are
// read (not the padding tail of the plane stride).
// ---------------------------------------------------------------------------
constexpr int kMaxStateThreads = 256;
constexpr int kMaxStateParts = 2048;  // partials per call (3 doubles each)

// the allocation-free variant's partials: one copy per GPU, 48 KiB
static __device__ __align__(256) double scratch[3 * kMaxStateParts];

__device__ inline void max_state_wg_reduce(double& mf, double& me, double& bad) {
  for (int o = 32; o > 0; o >>= 1) {
    mf = fmax(mf, __shfl_xor(mf, o));
    me = fmax(me, __shfl_xor(me, o));
    bad += __shfl_xor(bad, o);
  }
  __shared__ double red[3][kMaxStateThreads / 64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { red[0][w] = mf; red[1][w] = me; red[2][w] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kMaxStateThreads / 64; ++k) {
      mf = fmax(mf, red[0][k]);
      me = fmax(me, red[1][k]);
      bad += red[2][k];
    }
  }
}

template <typename V>
__device__ inline void max_state_plane(const V* p, long long nodes, double& m, double& bad) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += stride) {
    const double v = (double)p[i];
    if (isfinite(v)) m = fmax(m, fabs(v));
    else bad += 1.0;
  }
}

// F: 9 planes of S; eta: neta planes of T; plane stride P, nodes per plane
template <typename S, typename T>
__global__ __launch_bounds__(kMaxStateThreads) void k_max_state_part(
    const S* F, const T* eta, long long P, long long nodes, int neta, double* part) {
  double mf = 0.0, me = 0.0, bad = 0.0;
  for (int c = 0; c < 9; ++c) max_state_plane(F + c * P, nodes, mf, bad);
  for (int c = 0; c < neta; ++c) max_state_plane(eta + c * P, nodes, me, bad);
  max_state_wg_reduce(mf, me, bad);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = mf;
    part[3 * blockIdx.x + 1] = me;
    part[3 * blockIdx.x + 2] = bad;
  }
}

__global__ __launch_bounds__(kMaxStateThreads) void k_max_state_final(const double* part,
                                                                      int nparts, double* out) {
  double mf = 0.0, me = 0.0, bad = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    mf = fmax(mf, part[3 * i]);
    me = fmax(me, part[3 * i + 1]);
    bad += part[3 * i + 2];
  }
  max_state_wg_reduce(mf, me, bad);
  if (threadIdx.x == 0) { out[0] = mf; out[1] = me; out[2] = bad; }
}

// workgroups of the first pass: one per 8 * 256 nodes of a plane, at most
// kMaxStateParts (the grid-stride loop covers any block size)
static int max_state_parts(const wql_block* b) {
  const long long nodes = (long long)b->nx * b->ny * b->nz;
  const long long g = (nodes + 8LL * kMaxStateThreads - 1) / (8LL * kMaxStateThreads);
  return (int)std::max(1LL, std::min<long long>(g, kMaxStateParts));
}

// both passes on the block's stream with partials `part` (3 * nparts
// doubles) and the result in `out` (3 doubles, device)
static hipError_t launch_max_state(const wql_block* b, double* part, int nparts, double* out) {
  const long long nodes = (long long)b->nx * b->ny * b->nz;
  const int neta = 6 * b->d.nmech;
  const dim3 grid((unsigned)nparts), block(kMaxStateThreads);
  const void* F = b->F[b->cur];
  if (b->d.precision == WQL_FP64)
    hipLaunchKernelGGL((k_max_state_part<double, double>), grid, block, 0, b->stream,
                       (const double*)F, (const double*)b->eta, b->P, nodes, neta, part);
  else if (b->d.precision == WQL_FP32_MIXED)
    hipLaunchKernelGGL((k_max_state_part<double, float>), grid, block, 0, b->stream,
                       (const double*)F, (const float*)b->eta, b->P, nodes, neta, part);
  else
    hipLaunchKernelGGL((k_max_state_part<float, float>), grid, block, 0, b->stream,
                       (const float*)F, (const float*)b->eta, b->P, nodes, neta, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_max_state_final, dim3(1), block, 0, b->stream, (const double*)part, nparts,
                     out);
  return hipGetLastError();
}

// device address of `scratch` on each GPU (each holds its own copy), looked
// up once per device with that device current: the array's name is a
// host-side stand-in, never a device address
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

static int scratch_for_device(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(WQL_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(WQL_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return WQL_OK;
}

extern "C" int wql_block_max_state(wql_block* b, double* out) {
  if (!b || !out) return fail(WQL_EINVAL, "null argument");
  if (int rc = join(b)) return rc;
  const int nparts = max_state_parts(b);
  double* buf = nullptr;
  HIP_TRY(hipMalloc((void**)&buf, (size_t)(3 * nparts + 3) * sizeof(double)));
  hipError_t e = launch_max_state(b, buf, nparts, buf + 3 * nparts);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, buf + 3 * nparts, 3 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(WQL_EHIP, "max state: %s", hipGetErrorString(e));
  return WQL_OK;
}

extern "C" int wql_block_max_state_noalloc(wql_block* b, double* out_dev) {
  if (!b || !out_dev) return fail(WQL_EINVAL, "null argument");
  double* part = nullptr;
  if (int rc = scratch_for_device(b->device, &part)) return rc;
  const int nparts = max_state_parts(b);
  if (nparts > kMaxStateParts)
    return fail(WQL_EINVAL, "max state needs %d partials; the scratch array holds %d", nparts,
                kMaxStateParts);
  if (int rc = join(b)) return rc;
  HIP_TRY(launch_max_state(b, part, nparts, out_dev));
  return WQL_OK;
}

static int g_device = -1;

extern "C" const char* wql_last_error(void) { return g_err; }

extern "C" int wql_runtime_init(int32_t local_rank, int32_t* device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(WQL_ENODEV, "no GPU visible (hipGetDeviceCount: %s)",
                                              hipGetErrorString(e));
  const int dev = (local_rank < 0 ? 0 : local_rank) % n;
  HIP_TRY(hipSetDevice(dev));
  g_device = dev;
  // device address of the allocation-free entry points' scratch array on
  // this device (blocks created on another device look theirs up on first use)
  double* scratch_ptr = nullptr;
  if (int rc = scratch_for_device(dev, &scratch_ptr)) return rc;
  if (device) *device = dev;
  return WQL_OK;
}

extern "C" int wql_runtime_finalize(void) {
  if (g_device >= 0) HIP_TRY(hipDeviceSynchronize());
  g_device = -1;
  return WQL_OK;
}

extern "C" int wql_capacity_check(int64_t predicted) {
  size_t freeb = 0, total = 0;
  const char* env = getenv("WQL3D_GPU_MEMORY_BYTES");
  if (env && *env) {
    freeb = (size_t)strtoull(env, nullptr, 10);
  } else {
    HIP_TRY(hipMemGetInfo(&freeb, &total));
  }
  if (predicted < 0 || (uint64_t)predicted > (uint64_t)freeb)
    return fail(WQL_ENOMEM, "RUN-GPU-MEM-002: predicted %lld bytes exceed device capacity %llu",
                (long long)predicted, (unsigned long long)freeb);
  return WQL_OK;
}

// ---------------------------------------------------------------------------
static long long face_size(const wql_block* b, int f) {
  const int axis = f >> 1;
  return axis == 0 ? (long long)b->ny * b->nz : (axis == 1 ? (long long)b->nx * b->nz
                                                           : (long long)b->nx * b->ny);
}

static int validate(const wql_block_desc* d) {
  if (!d) return fail(WQL_EINVAL, "null descriptor");
  const int n[3] = {d->nq, d->nr, d->ns};
  for (int k = 0; k < 3; ++k)
    if (n[k] < 12) return fail(WQL_EINVAL, "axis %d has %d nodes; SBP-6 needs >= 12", k, n[k]);
  // kernels address planes with 32-bit byte offsets: 8 * plane stride < 2^32
  if ((long long)n[0] * n[1] * n[2] + kPlaneAlign >= (1LL << 29))
    return fail(WQL_EINVAL, "block of %lld nodes exceeds 2^29 nodes per block",
                (long long)n[0] * n[1] * n[2]);
  if (d->mesh != WQL_CARTESIAN && d->mesh != WQL_CURVILINEAR)
    return fail(WQL_EINVAL, "invalid mesh %d", d->mesh);
  if (d->precision != WQL_FP64 && d->precision != WQL_FP32_MIXED &&
      d->precision != WQL_FP32_KAHAN)
    return fail(WQL_EINVAL, "invalid precision %d", d->precision);
  if (d->nmech != 0 && d->nmech != 4 && d->nmech != 8)
    return fail(WQL_EINVAL, "nmech must be 0, 4 or 8 (got %d)", d->nmech);
  for (int f = 0; f < 6; ++f) {
    if (d->bc[f] > 3) return fail(WQL_EINVAL, "invalid bc code %d on face %d", d->bc[f], f);
    if (d->pml[f]) {
      if (d->npml[f] < 1 || d->npml[f] > n[f >> 1])
        return fail(WQL_EINVAL, "invalid npml %d on face %d", d->npml[f], f);
    }
  }
  for (int a = 0; a < 3; ++a) {
    const int lo = d->pml[2 * a] ? d->npml[2 * a] : 0, hi = d->pml[2 * a + 1] ? d->npml[2 * a + 1] : 0;
    if (lo + hi > n[a]) return fail(WQL_EINVAL, "PML slabs overlap on axis %d", a);
  }
  if (d->nmech)
    for (int i = 0; i < d->nmech; ++i)
      if (!(d->tau[i] > 0.0)) return fail(WQL_EINVAL, "tau[%d] must be positive", i);
  if (d->hq <= 0 || d->hr <= 0 || d->hs <= 0) return fail(WQL_EINVAL, "h must be positive");
  if (d->sat_form != WQL_SAT_BC && d->sat_form != WQL_SAT_ALPHA_THETA)
    return fail(WQL_EINVAL, "invalid sat_form %d", d->sat_form);
  if (d->sat_form == WQL_SAT_ALPHA_THETA)
    for (int f = 0; f < 6; ++f)
      if (d->pml[f]) return fail(WQL_EINVAL, "the alpha/theta SAT form takes no PML faces (face %d)", f);
  return WQL_OK;
}

static int dalloc(wql_block* b, void** p, size_t bytes) {
  if (bytes == 0) { *p = nullptr; return WQL_OK; }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(WQL_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b->bytes += (long long)bytes;
  return WQL_OK;
}

// live blocks, for wql_block_find (host F pointer -> block)
static std::mutex g_blocks_mu;
static std::vector<wql_block*> g_blocks;

static void register_block(wql_block* b) {
  std::lock_guard<std::mutex> lock(g_blocks_mu);
  g_blocks.push_back(b);
}

extern "C" int wql_block_find(const double* host_F, wql_block** out) {
  if (!out) return fail(WQL_EINVAL, "null output pointer");
  *out = nullptr;
  std::lock_guard<std::mutex> lock(g_blocks_mu);
  for (wql_block* b : g_blocks)
    if (host_F && b->d.F == host_F) {
      *out = b;
      return WQL_OK;
    }
  return fail(WQL_EINVAL, "no block was created with F at %p", (const void*)host_F);
}

extern "C" void wql_block_destroy(wql_block* b) {
  if (!b) return;
  {
    std::lock_guard<std::mutex> lock(g_blocks_mu);
    g_blocks.erase(std::remove(g_blocks.begin(), g_blocks.end(), b), g_blocks.end());
  }
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  (void)hipStreamSynchronize(b->stream);
  if (b->iface) {  // the interface outlives the block: it refuses further stages
    if (b->iface->L == b) b->iface->L = nullptr;
    if (b->iface->R == b) b->iface->R = nullptr;
  }
  void* ptrs[] = {b->arena, b->rec_idx, b->seis, b->X, b->adj_Fb, b->adj_Eb, b->adj_DFb,
                  b->adj_DEb, b->adj_g, b->adj_Z, b->slot_F, b->slot_E};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 2; ++k) {
    if (b->stage_buf[k]) (void)hipHostFree(b->stage_buf[k]);
    if (b->stage_ev[k]) (void)hipEventDestroy(b->stage_ev[k]);
  }
  for (int f = 0; f < 6; ++f) {
    if (b->pQ[f]) (void)hipFree(b->pQ[f]);
    if (b->pDQ[f]) (void)hipFree(b->pDQ[f]);
    if (b->face[f]) (void)hipFree(b->face[f]);
    free(b->face64[f]);
  }
  for (int i = 0; i < b->nev_alloc; ++i) (void)hipEventDestroy(b->ev[i]);
  free(b->ev);
  free(b->slot_mz);
  for (int k = 0; k < kMaxChunks; ++k) {
    if (b->ev_sten[k]) (void)hipEventDestroy(b->ev_sten[k]);
    if (b->ev_mem[k]) (void)hipEventDestroy(b->ev_mem[k]);
  }
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  (void)hipStreamDestroy(b->stream);
  free(b);
}

// time the memory-variable kernel on a candidate arena (zero-filled first)
static int arena_probe_ms(wql_block* b, void* arena, size_t bytes, float* ms);

extern "C" int wql_block_create(const wql_block_desc* d, wql_block** out) {
  if (!out) return fail(WQL_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = validate(d);
  if (rc) return rc;
  wql_block* b = (wql_block*)calloc(1, sizeof(wql_block));
  if (!b) return fail(WQL_ENOMEM, "host allocation failed");
  b->d = *d;
  if (hipGetDevice(&b->device) != hipSuccess) b->device = 0;
  b->nx = d->nq; b->ny = d->nr; b->nz = d->ns;
  b->P = ((long long)b->nx * b->ny * b->nz + kPlaneAlign - 1) / kPlaneAlign * kPlaneAlign;
  b->esz = d->precision == WQL_FP64 ? sizeof(double) : sizeof(float);
  b->ssz = d->precision == WQL_FP32_KAHAN ? sizeof(float) : sizeof(double);
  // launch shape (env WQL3D_NODE_TUNING = "nontemporal,xcd_remap,band_rows,
  // split,interior,zchunk"): xcd_remap 0 = round-robin, 1 = contiguous,
  // G >= 2 = groups of G; band_rows 0 = whole y extent; split = memory
  // variables of the shell in their own kernel; interior = tile of the
  // strict-interior kernel (0 off, 1 64x8, 2 32x16, 3 64x4); zchunk = z
  // levels per interior workgroup
  b->tune_nt = 0;
  // workgroup order over the XCDs: groups of 16 consecutive workgroups per
  // XCD (xcd_logical).  At 385^3 Q8 FP32-mixed on one allocation the
  // interior kernel takes 2.74-2.81 ms against 2.96-3.00 with the hardware's
  // round-robin order and the shell 3.22-3.30 against 3.34-3.35 (groups of
  // 4-64 all gain; FP64 is unchanged within 0.1 ms); the contiguous order
  // (1) is slower (shell 4.05 ms)
  b->tune_xcd = 16;
  b->tune_band = 0;
  b->tune_split = 1;
  b->tune_zm = 1;
  // 24 z levels per interior workgroup: on one allocation at 385^3, 3.52 ->
  // 3.46 / 3.34 -> 3.30 ms (FP32-mixed) and 5.16 -> 5.12 ms (FP64) against 32
  b->tune_zchunk = 24;
  if (const char* tv = getenv("WQL3D_NODE_TUNING")) {
    int v0 = 0, v1 = 16, v2 = 0, v3 = 1, v4 = 1, v5 = 24;
    if (sscanf(tv, "%d,%d,%d,%d,%d,%d", &v0, &v1, &v2, &v3, &v4, &v5) >= 3) {
      b->tune_nt = v0 != 0;
      b->tune_xcd = v1 > 0 ? v1 : 0;
      b->tune_band = v2 > 0 ? v2 : 0;
      b->tune_split = v3 != 0;
      b->tune_zm = (v4 >= 0 && v4 <= 3) ? v4 : 1;
      b->tune_zchunk = v5 > 0 ? v5 : 24;
    }
  }
  // non-temporal vector streams in the memory-variable kernel: 9.67 -> 9.09 ms
  // at 385^3 Q8 FP32-mixed (every plane is read or written once per stage)
  b->tune_mem_nt = 1;
  if (const char* nv = getenv("WQL3D_MEM_NT")) b->tune_mem_nt = atoi(nv) != 0;
  // the same grouped workgroup order for the memory-variable kernel (no
  // stencil reuse, but the ~180 plane streams in flight then come from one
  // compact node range): at 385^3 Q8 on arena-searched blocks, FP32-mixed
  // 7.78-7.82 -> 7.61-7.62 ms, FP64 12.77-12.97 -> 12.39-12.51 ms (groups
  // of 64 the same, 256 and the contiguous order no gain)
  b->tune_mem_xcd = 16;
  // and for the adjoint's per-node kernels: reverse step at 129^3 Q4 8.64-8.74
  // -> 8.41-8.45 ms FP64, 4.67-4.68 -> 4.46 ms FP32-mixed
  b->tune_adj_xcd = 16;
  if (const char* av = getenv("WQL3D_ADJ_XCD")) b->tune_adj_xcd = atoi(av) > 0 ? atoi(av) : 0;
  if (const char* xv = getenv("WQL3D_MEM_XCD")) b->tune_mem_xcd = atoi(xv) > 0 ? atoi(xv) : 0;  // 0 = round-robin
  b->tune_mem_npt = 2;
  if (const char* mv = getenv("WQL3D_MEM_NPT")) {
    const int v = atoi(mv);
    b->tune_mem_npt = v == 1 ? 1 : (v == 4 ? 4 : 2);
  }
  b->nchunks = kDefaultChunks;
  if (const char* cv = getenv("WQL3D_STAGE_CHUNKS")) {
    const int v = atoi(cv);
    b->nchunks = v < 1 ? 1 : (v > kMaxChunks ? kMaxChunks : v);
  }
  // a chunk must hold at least 3 z levels (the stencil's z halo)
  while (b->nchunks > 1 && b->nz / b->nchunks < 3) --b->nchunks;
  hipError_t e = hipStreamCreate(&b->stream);  // default priority
  if (e != hipSuccess) { free(b); return fail(WQL_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  e = hipStreamCreate(&b->stream2);  // default priority
  for (int k = 0; k < kMaxChunks && e == hipSuccess; ++k) {
    e = hipEventCreateWithFlags(&b->ev_sten[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_mem[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    wql_block_destroy(b);
    return fail(WQL_EHIP, "stream / event creation: %s", hipGetErrorString(e));
  }
  const long long P = b->P;
  const int ngeo = d->mesh == WQL_CURVILINEAR ? 10 : 3;
#define ALLOC(ptr, bytes)                                  \
  do {                                                     \
    int r_ = dalloc(b, (void**)&(ptr), (size_t)(bytes));   \
    if (r_) { wql_block_destroy(b); return r_; }           \
  } while (0)
  {
    // all per-node plane arrays in one arena: array k starts on a 2 MiB
    // boundary plus (k * stagger) mod 2 MiB bytes (env WQL3D_ARENA_STAGGER)
    struct Part { void** ptr; long long bytes; };
    const long long nm = d->nmech;
    const Part parts[] = {
        {&b->F[0], 9 * P * (long long)b->ssz}, {&b->F[1], 9 * P * (long long)b->ssz},
        {&b->Fo[0], d->precision == WQL_FP32_MIXED ? 9 * P * 4 : 0},
        {&b->Fo[1], d->precision == WQL_FP32_MIXED ? 9 * P * 4 : 0},
        {&b->DF, 9 * P * (long long)b->ssz},
        {&b->Fc, d->precision == WQL_FP32_KAHAN ? 9 * P * 4 : 0},
        {&b->G, ngeo * P * (long long)b->esz}, {&b->M, 3 * P * (long long)b->esz},
        {&b->eta, 6 * nm * P * (long long)b->esz}, {&b->Deta, 6 * nm * P * (long long)b->esz},
        {&b->Qs, nm ? P * (long long)b->esz : 0}, {&b->Qp, nm ? P * (long long)b->esz : 0},
        {&b->W, nm ? 12 * P * (long long)b->esz : 0}};
    const long long kPage = 2LL << 20;
    long long stagger = 0;
    if (const char* sv = getenv("WQL3D_ARENA_STAGGER")) stagger = atoll(sv);
    long long off[sizeof parts / sizeof parts[0]], end = 0;
    int k = 0;
    for (const Part& p : parts) {
      off[k] = (end + kPage - 1) / kPage * kPage + ((k * stagger) % kPage + kPage) % kPage / 256 * 256;
      end = off[k] + p.bytes;
      ++k;
    }
    auto carve = [&](void* base) {
      int i = 0;
      for (const Part& p : parts) {
        *p.ptr = p.bytes ? (void*)((char*)base + off[i]) : nullptr;
        ++i;
      }
    };
    // The memory-variable kernel streams ~100 planes at once, and its speed
    // depends on where the arena lands in HBM: at 385^3 Q8 FP32-mixed one
    // allocation runs it in 8.1-8.2 ms and another in 9.4-9.5 ms (FP64
    // 13.4-13.9 against 16.4 ms), the same for the whole life of the
    // allocation, while the stencil kernels do not vary.  For a large
    // anelastic block, arenas are allocated one after another (each held, so
    // the next lands elsewhere) while the device has room, the kernel is timed
    // on each (zero-filled, so it computes zeros), and the fastest is kept.
    // The search stops at `tries` arenas (env WQL3D_ARENA_TRIES, 1 = take the
    // first) or once both modes have been seen (slowest >= 1.07 x fastest).
    int tries = 6;
    if (const char* tv = getenv("WQL3D_ARENA_TRIES")) tries = std::max(1, atoi(tv));
    if (d->nmech == 0 || !b->tune_split || end < (1LL << 32)) tries = 1;
    std::vector<void*> cand;
    int best = -1;
    float best_ms = 0.0f, worst_ms = 0.0f;
    for (int t = 0; t < tries; ++t) {
      if (t > 0) {
        size_t freeb = 0, totb = 0;
        if (hipMemGetInfo(&freeb, &totb) != hipSuccess ||
            (long long)freeb < end + end / 8 + (4LL << 30))
          break;
      }
      void* p = nullptr;
      if (hipMalloc(&p, (size_t)end) != hipSuccess) {
        (void)hipGetLastError();
        break;
      }
      cand.push_back(p);
      if (tries == 1) { best = 0; break; }
      carve(p);
      float ms = 0.0f;
      if (arena_probe_ms(b, p, (size_t)end, &ms) != WQL_OK) { best = (int)cand.size() - 1; break; }
      if (getenv("WQL3D_ARENA_VERBOSE"))
        fprintf(stderr, "[wql3d] arena candidate %d: memory kernel %.3f ms\n", t, ms);
      if (best < 0 || ms < best_ms) { best = (int)cand.size() - 1; best_ms = ms; }
      if (ms > worst_ms) worst_ms = ms;
      if (worst_ms >= 1.07f * best_ms) break;
    }
    if (best < 0) {
      for (void* p : cand) (void)hipFree(p);
      wql_block_destroy(b);
      return fail(WQL_ENOMEM, "hipMalloc(%lld) failed for the block's arena", end);
    }
    for (int i = 0; i < (int)cand.size(); ++i)
      if (i != best) (void)hipFree(cand[i]);
    b->arena = cand[best];
    b->bytes += end;
    carve(b->arena);
  }
  const int n[3] = {b->nx, b->ny, b->nz};
  for (int f = 0; f < 6; ++f) {
    if (d->pml[f]) {
      const int ax = f >> 1;
      for (int k = 0; k < 3; ++k) { b->slab_lo[f][k] = 0; b->slab_n[f][k] = n[k]; }
      b->slab_n[f][ax] = d->npml[f];
      b->slab_lo[f][ax] = (f & 1) ? n[ax] - d->npml[f] : 0;
      b->Ps[f] = (long long)b->slab_n[f][0] * b->slab_n[f][1] * b->slab_n[f][2];
      ALLOC(b->pQ[f], 9 * b->Ps[f] * b->esz);
      ALLOC(b->pDQ[f], 9 * b->Ps[f] * b->esz);
    }
    // face frames on boundary faces, and on bc 0 faces given frames (interface faces)
    if (d->bc[f] > 0 || (d->bc[f] == 0 && d->n_l[f] && d->n_m[f] && d->n_n[f])) {
      b->NF[f] = face_size(b, f);
      ALLOC(b->face[f], 13 * b->NF[f] * b->esz);
    }
  }
#undef ALLOC
  *out = b;
  register_block(b);
  return wql_block_init_state(b);
}

extern "C" int64_t wql_block_device_bytes(const wql_block* b) { return b ? b->bytes : 0; }

extern "C" int wql_block_sync(wql_block* b) {
  if (!b) return fail(WQL_EINVAL, "null block");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return WQL_OK;
}

extern "C" void* wql_block_stream(wql_block* b) { return b ? (void*)b->stream : nullptr; }

// ---------------------------------------------------------------------------
// upload / download of one 3-D component between a Fortran-bounded double
// array and a dense device plane of T.  (x0,y0,z0): 0-based origin of the
// node range inside the block, (mx,my,mz): its extents.
// ---------------------------------------------------------------------------
struct View {
  const double* base;  // element (lo0, lo1, lo2) of component 0
  wql_box box;
  long long sx, sy, sz, sc;
};

static View make_view(const double* base, const wql_box& box) {
  View v;
  v.base = base;
  v.box = box;
  v.sx = 1;
  v.sy = box_extent(box, 0);
  v.sz = v.sy * box_extent(box, 1);
  v.sc = v.sz * box_extent(box, 2);
  return v;
}

// offset of 1-based node (x1,y1,z1) component c in the view
static long long view_off(const View& v, long long x1, long long y1, long long z1, int c) {
  return (x1 - v.box.lo[0]) * v.sx + (y1 - v.box.lo[1]) * v.sy + (z1 - v.box.lo[2]) * v.sz +
         (long long)c * v.sc;
}

static int check_view(const View& v, const int lo1[3], const int n[3], const char* name) {
  for (int k = 0; k < 3; ++k)
    if (lo1[k] < v.box.lo[k] || lo1[k] + n[k] - 1 > v.box.hi[k])
      return fail(WQL_EINVAL, "array %s bounds [%d,%d] on dim %d do not cover nodes [%d,%d]", name,
                  v.box.lo[k], v.box.hi[k], k + 1, lo1[k], lo1[k] + n[k] - 1);
  return WQL_OK;
}

// Run body(z0, z1) over the z levels [0, nz) on up to 16 host threads, one
// per 256 Ki values at least (a small component is packed on this thread:
// starting threads would cost more than the copy).
template <typename Fn>
static void host_parallel_z(long long nz, long long values, Fn&& body) {
  const unsigned hw = std::thread::hardware_concurrency();
  long long nt = std::min<long long>(nz, std::max(1u, std::min(hw, 16u)));
  nt = std::min<long long>(nt, std::max(1LL, values >> 18));
  if (nt <= 1) { body(0LL, nz); return; }
  std::vector<std::thread> th;
  long long started = 0;
  try {
    th.reserve((size_t)nt);
    for (; started < nt; ++started) {
      const long long t = started;
      th.emplace_back([&, t] { body(nz * t / nt, nz * (t + 1) / nt); });
    }
  } catch (...) {
    // thread creation failed (std::system_error / bad_alloc): the threads
    // already running keep their ranges, this thread does the rest
  }
  if (started < nt) body(nz * started / nt, nz);
  for (auto& x : th) x.join();
}

// pinned staging buffer for one plane of `bytes` (bytes <= one plane of
// doubles), waiting until its previous copy has finished
static int staging(wql_block* b, size_t bytes, void** out) {
  if (!b->stage_buf[0]) {
    const size_t plane = (size_t)b->nx * b->ny * b->nz * sizeof(double);
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(hipHostMalloc(&b->stage_buf[k], plane, hipHostMallocDefault));
      HIP_TRY(hipEventCreateWithFlags(&b->stage_ev[k], hipEventDisableTiming));
      HIP_TRY(hipEventRecord(b->stage_ev[k], b->stream));
    }
    b->stage_bytes = plane;
    b->stage_k = 0;
  }
  if (bytes > b->stage_bytes) return fail(WQL_EINVAL, "transfer larger than a plane");
  const int k = b->stage_k;
  HIP_TRY(hipEventSynchronize(b->stage_ev[k]));
  *out = b->stage_buf[k];
  return WQL_OK;
}

// the copy out of (or into) the current staging buffer has been issued
static int staging_issued(wql_block* b) {
  HIP_TRY(hipEventRecord(b->stage_ev[b->stage_k], b->stream));
  b->stage_k ^= 1;
  return WQL_OK;
}

template <typename T>
static int put_comp(wql_block* b, T* dst, const View& v, int c, const int lo1[3], const int n[3],
                    const char* name) {
  if (!v.base) return fail(WQL_EINVAL, "array %s is null", name);
  int rc = check_view(v, lo1, n, name);
  if (rc) return rc;
  const long long off0 = view_off(v, lo1[0], lo1[1], lo1[2], c);
  const long long cnt = (long long)n[0] * n[1] * n[2];
  if (b->d.arrays_on_device) {
    HIP_TRY(launch_gather<T>(dst, v.base + off0, n[0], n[1], n[2], 1, v.sx, v.sy, v.sz, 0, 0,
                             b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return WQL_OK;
  }
  // pack (and convert) into pinned memory on host threads, then copy; the
  // copy overlaps the packing of the next component
  void* buf = nullptr;
  if ((rc = staging(b, (size_t)cnt * sizeof(T), &buf))) return rc;
  T* h = (T*)buf;
  host_parallel_z(n[2], cnt, [&](long long z0, long long z1) {
    for (long long z = z0; z < z1; ++z)
      for (long long y = 0; y < n[1]; ++y) {
        const double* src = v.base + off0 + y * v.sy + z * v.sz;
        T* o = h + (size_t)(n[0] * (y + (long long)n[1] * z));
        for (long long x = 0; x < n[0]; ++x) o[x] = (T)src[x];
      }
  });
  HIP_TRY(hipMemcpyAsync(dst, h, cnt * sizeof(T), hipMemcpyHostToDevice, b->stream));
  return staging_issued(b);
}

template <typename T>
static int get_comp(wql_block* b, const T* src, const View& v, int c, const int lo1[3],
                    const int n[3], const char* name) {
  if (!v.base) return fail(WQL_EINVAL, "array %s is null", name);
  int rc = check_view(v, lo1, n, name);
  if (rc) return rc;
  const long long off0 = view_off(v, lo1[0], lo1[1], lo1[2], c);
  const long long cnt = (long long)n[0] * n[1] * n[2];
  double* base = const_cast<double*>(v.base);
  if (b->d.arrays_on_device) {
    HIP_TRY(launch_scatter<T>(base + off0, src, n[0], n[1], n[2], 1, v.sx, v.sy, v.sz, 0, 0,
                              b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return WQL_OK;
  }
  void* buf = nullptr;
  if ((rc = staging(b, (size_t)cnt * sizeof(T), &buf))) return rc;
  T* h = (T*)buf;
  HIP_TRY(hipMemcpyAsync(h, src, cnt * sizeof(T), hipMemcpyDeviceToHost, b->stream));
  if ((rc = staging_issued(b))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  host_parallel_z(n[2], cnt, [&](long long z0, long long z1) {
    for (long long z = z0; z < z1; ++z)
      for (long long y = 0; y < n[1]; ++y) {
        double* o = base + off0 + y * v.sy + z * v.sz;
        const T* s = h + (size_t)(n[0] * (y + (long long)n[1] * z));
        for (long long x = 0; x < n[0]; ++x) o[x] = (double)s[x];
      }
  });
  return WQL_OK;
}

// face frame + |grad xi| + metric column planes of face f (unit_normals.f90:
// 8-168 results, norm as in BoundaryConditions.f90:114-118, the metric column
// as read by Couple_Interface_x, CouplingForcing.f90:228-260)
template <typename T>
static int put_face(wql_block* b, int f) {
  const wql_block_desc& d = b->d;
  const int axis = f >> 1;
  const int n1 = axis == 0 ? b->ny : b->nx, n2 = axis == 2 ? b->ny : b->nz;
  const long long NF = (long long)n1 * n2;
  const wql_box& fb = d.face_box[f];
  if (!d.n_l[f] || !d.n_m[f] || !d.n_n[f]) return fail(WQL_EINVAL, "face %d frame arrays null", f);
  if (fb.lo[0] > 1 || fb.hi[0] < n1 || fb.lo[1] > 1 || fb.hi[1] < n2)
    return fail(WQL_EINVAL, "face %d frame bounds do not cover the face", f);
  const long long e0 = box_extent(fb, 0), e1 = box_extent(fb, 1);
  const int fixed1 = (f & 1) ? (axis == 0 ? b->nx : (axis == 1 ? b->ny : b->nz)) : 1;
  std::vector<double> frame(9 * NF), met(3 * NF);
  // frames
  const double* fr[3] = {d.n_l[f], d.n_m[f], d.n_n[f]};
  std::vector<double> tmp;
  for (int k = 0; k < 3; ++k) {
    const double* src = fr[k];
    if (d.arrays_on_device) {
      tmp.resize(3 * e0 * e1);
      HIP_TRY(hipMemcpy(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
      src = tmp.data();
    }
    for (int c = 0; c < 3; ++c)
      for (long long j = 0; j < n2; ++j)
        for (long long i = 0; i < n1; ++i)
          frame[(3 * k + c) * NF + i + n1 * j] =
              src[(i + 1 - fb.lo[0]) + e0 * ((j + 1 - fb.lo[1]) + e1 * c)];
  }
  // metric column `axis` of metricx/y/z at the face nodes
  const double* mets[3] = {d.metricx, d.metricy, d.metricz};
  const View body0 = make_view(nullptr, d.body);
  const long long fx1 = axis == 0 ? fixed1 : 1, fy1 = axis == 1 ? fixed1 : 1, fz1 = axis == 2 ? fixed1 : 1;
  const int dims[3] = {axis == 0 ? 1 : b->nx, axis == 1 ? 1 : b->ny, axis == 2 ? 1 : b->nz};
  for (int m = 0; m < 3; ++m) {
    if (!mets[m]) return fail(WQL_EINVAL, "metric array %d null", m);
    const long long off0 = view_off(body0, fx1, fy1, fz1, axis);
    if (d.arrays_on_device) {
      double* tmpd = nullptr;
      HIP_TRY(hipMalloc((void**)&tmpd, NF * sizeof(double)));
      hipError_t e = launch_gather<double>(tmpd, mets[m] + off0, dims[0], dims[1], dims[2], 1,
                                           body0.sx, body0.sy, body0.sz, 0, 0, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e == hipSuccess)
        e = hipMemcpy(met.data() + m * NF, tmpd, NF * sizeof(double), hipMemcpyDeviceToHost);
      (void)hipFree(tmpd);
      if (e != hipSuccess) return fail(WQL_EHIP, "face metric gather: %s", hipGetErrorString(e));
    } else {
      for (long long z = 0; z < dims[2]; ++z)
        for (long long y = 0; y < dims[1]; ++y)
          for (long long x = 0; x < dims[0]; ++x)
            met[m * NF + x + dims[0] * (y + (long long)dims[1] * z)] =
                mets[m][off0 + x * body0.sx + y * body0.sy + z * body0.sz];
    }
  }
  if (!b->face64[f]) {
    b->face64[f] = (double*)malloc((size_t)(13 * NF) * sizeof(double));
    if (!b->face64[f]) return fail(WQL_ENOMEM, "host allocation failed");
  }
  double* o64 = b->face64[f];
  for (long long k = 0; k < 9 * NF; ++k) o64[k] = frame[k];
  for (long long k = 0; k < NF; ++k) {
    const double qx = met[k], qy = met[NF + k], qz = met[2 * NF + k];
    o64[9 * NF + k] = sqrt(qx * qx + qy * qy + qz * qz);
    o64[10 * NF + k] = qx;
    o64[11 * NF + k] = qy;
    o64[12 * NF + k] = qz;
  }
  std::vector<T> out(13 * NF);
  for (long long k = 0; k < 13 * NF; ++k) out[k] = (T)o64[k];
  HIP_TRY(hipMemcpy(b->face[f], out.data(), out.size() * sizeof(T), hipMemcpyHostToDevice));
  return WQL_OK;
}

template <typename T>
static int transfer(wql_block* b, uint32_t mask, bool up) {
  const wql_block_desc& d = b->d;
  const long long P = b->P;
  const int lo1[3] = {1, 1, 1};
  const int n[3] = {b->nx, b->ny, b->nz};
  const View body = make_view(nullptr, d.body);
  int rc = 0;
  auto xfer = [&](void* dev, const double* host, int c, const char* nm, const int* l,
                  const int* nn, const View& vtemplate) -> int {
    View v = vtemplate;
    v.base = host;
    return up ? put_comp<T>(b, (T*)dev, v, c, l, nn, nm) : get_comp<T>(b, (const T*)dev, v, c, l, nn, nm);
  };
  // the RK state: double planes, or float planes in FP32-Kahan
  const bool kahan = d.precision == WQL_FP32_KAHAN;
  auto xfer_state = [&](void* base, const double* host, int c, const char* nm) -> int {
    View v = body;
    v.base = host;
    if (kahan) {
      float* dev = (float*)base + c * P;
      return up ? put_comp<float>(b, dev, v, c, lo1, n, nm) : get_comp<float>(b, dev, v, c, lo1, n, nm);
    }
    double* dev = (double*)base + c * P;
    return up ? put_comp<double>(b, dev, v, c, lo1, n, nm) : get_comp<double>(b, dev, v, c, lo1, n, nm);
  };
  if (mask & WQL_F) {
    if (up) b->fo_valid = false;
    for (int c = 0; c < 9 && !rc; ++c) rc = xfer_state(b->F[b->cur], d.F, c, "F");
    if (!rc && up && kahan) HIP_TRY(launch_fill<float>((float*)b->Fc, 9 * P, 0.0, b->stream));
  }
  if (mask & WQL_DF)
    for (int c = 0; c < 9 && !rc; ++c) rc = xfer_state(b->DF, d.DF, c, "DF");
  if (rc) return rc;
  if (mask & WQL_GEOMETRY) {
    T* G = (T*)b->G;
    if (d.mesh == WQL_CURVILINEAR) {
      const double* m3[3] = {d.metricx, d.metricy, d.metricz};
      for (int m = 0; m < 3 && !rc; ++m)
        for (int c = 0; c < 3 && !rc; ++c)
          rc = xfer(G + (3 * m + c) * P, m3[m], c, "metric", lo1, n, body);
      if (!rc) rc = xfer(G + 9 * P, d.J, 0, "J", lo1, n, body);
    } else {
      rc = xfer(G, d.metricx, 0, "metricx", lo1, n, body);
      if (!rc) rc = xfer(G + P, d.metricy, 1, "metricy", lo1, n, body);
      if (!rc) rc = xfer(G + 2 * P, d.metricz, 2, "metricz", lo1, n, body);
    }
    if (rc) return rc;
  }
  if (mask & WQL_MATERIAL)
    for (int c = 0; c < 3 && !rc; ++c) rc = xfer((T*)b->M + c * P, d.M, c, "M", lo1, n, body);
  if (rc) return rc;
  if (d.nmech && (mask & WQL_MEMORY)) {
    if (up) b->mech_zero = 0;  // uploaded memory variables may be non-zero
    for (int c6 = 0; c6 < 6 && !rc; ++c6)
      for (int i = 0; i < d.nmech && !rc; ++i) {
        rc = xfer((T*)b->eta + (c6 * d.nmech + i) * P, d.eta[c6], i, "eta", lo1, n, body);
        if (!rc) rc = xfer((T*)b->Deta + (c6 * d.nmech + i) * P, d.Deta[c6], i, "Deta", lo1, n, body);
      }
    if (rc) return rc;
  }
  if (d.nmech && (mask & WQL_QINV)) {
    rc = xfer(b->Qs, d.Qs_inv, 0, "Qs_inv", lo1, n, body);
    if (!rc) rc = xfer(b->Qp, d.Qp_inv, 0, "Qp_inv", lo1, n, body);
    if (rc) return rc;
  }
  if (mask & WQL_PML) {
    for (int f = 0; f < 6 && !rc; ++f) {
      if (!d.pml[f]) continue;
      const View pv = make_view(nullptr, d.pml_box[f]);
      int sl[3], sn[3];
      for (int k = 0; k < 3; ++k) { sl[k] = b->slab_lo[f][k] + 1; sn[k] = b->slab_n[f][k]; }
      for (int c = 0; c < 9 && !rc; ++c) {
        rc = xfer((T*)b->pQ[f] + c * b->Ps[f], d.pml_Q[f], c, "pml_Q", sl, sn, pv);
        if (!rc) rc = xfer((T*)b->pDQ[f] + c * b->Ps[f], d.pml_DQ[f], c, "pml_DQ", sl, sn, pv);
      }
    }
    if (rc) return rc;
  }
  if (up && (mask & WQL_FACES))
    for (int f = 0; f < 6 && !rc; ++f)
      if (b->face[f]) rc = put_face<T>(b, f);
  return rc;
}

static int iface_fill_faces(wql_iface* I);

extern "C" int wql_block_upload(wql_block* b, uint32_t mask) {
  if (!b) return fail(WQL_EINVAL, "null block");
  int rc = b->d.precision == WQL_FP64 ? transfer<double>(b, mask, true)
                                      : transfer<float>(b, mask, true);
  // the data is on the device when the call returns (other streams, e.g. an
  // interface's, may read it next)
  HIP_TRY(hipStreamSynchronize(b->stream));
  // a coupled block's new frames or material reach the interface's FP64 copies
  if (!rc && b->iface && b->iface->L && b->iface->R && (mask & (WQL_FACES | WQL_MATERIAL)))
    rc = iface_fill_faces(b->iface);
  return rc;
}

extern "C" int wql_block_download(wql_block* b, uint32_t mask) {
  if (!b) return fail(WQL_EINVAL, "null block");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return b->d.precision == WQL_FP64 ? transfer<double>(b, mask & ~WQL_FACES, false)
                                    : transfer<float>(b, mask & ~WQL_FACES, false);
}

template <typename T>
static int init_state_t(wql_block* b) {
  const long long P = b->P;
  const int nm = b->d.nmech;
  if (b->d.precision == WQL_FP32_KAHAN) {
    HIP_TRY(launch_fill<float>((float*)b->F[0], 9 * P, 0.0, b->stream));
    HIP_TRY(launch_fill<float>((float*)b->F[1], 9 * P, 0.0, b->stream));
    HIP_TRY(launch_fill<float>((float*)b->Fc, 9 * P, 0.0, b->stream));
    // 1e40 (fields.f90:39) is +inf in float; the A(1) = 0 scaling maps it to 0
    HIP_TRY(launch_fill<float>((float*)b->DF, 9 * P, 1.0e40, b->stream));
  } else {
    HIP_TRY(launch_fill<double>((double*)b->F[0], 9 * P, 0.0, b->stream));
    HIP_TRY(launch_fill<double>((double*)b->F[1], 9 * P, 0.0, b->stream));
    HIP_TRY(launch_fill<double>((double*)b->DF, 9 * P, 1.0e40, b->stream));   // fields.f90:39
  }
  if (nm) {
    HIP_TRY(launch_fill<T>((T*)b->eta, 6LL * nm * P, 0.0, b->stream));  // material.f90:389-404
    HIP_TRY(launch_fill<T>((T*)b->Deta, 6LL * nm * P, 0.0, b->stream));
  }
  for (int f = 0; f < 6; ++f)
    if (b->pQ[f]) {
      HIP_TRY(launch_fill<T>((T*)b->pQ[f], 9 * b->Ps[f], 0.0, b->stream));  // pml.f90:45-46
      // 1e40 in double; the fp32 mirror holds +inf, which the A(1)=0 scaling maps to 0
      HIP_TRY(launch_fill<T>((T*)b->pDQ[f], 9 * b->Ps[f], 1.0e40, b->stream));
    }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return WQL_OK;
}

extern "C" int wql_block_init_state(wql_block* b) {
  if (!b) return fail(WQL_EINVAL, "null block");
  b->cur = 0;
  b->fo_valid = false;
  b->mech_zero = (1 << b->d.nmech) - 1;  // eta = Deta = 0 after the fill below
  b->slot = 0;
  return b->d.precision == WQL_FP64 ? init_state_t<double>(b) : init_state_t<float>(b);
}

// ---------------------------------------------------------------------------
extern "C" int wql_block_set_sources(wql_block* b, int32_t n, const wql_source* src) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (n < 0 || n > kMaxSources) return fail(WQL_EINVAL, "at most %d sources", kMaxSources);
  for (int i = 0; i < n; ++i) {
    const wql_source& s = src[i];
    if (s.order != 0 && s.order != 2 && s.order != 4)
      return fail(WQL_EINVAL, "source order must be 2 or 4, or 0 (mollified)");
    if (s.order == 0 && !b->X)
      return fail(WQL_ESTATE, "a mollified source needs the node coordinates (wql_block_set_coordinates)");
    if (s.type < 0 || s.type > 2) return fail(WQL_EINVAL, "invalid source type %d", s.type);
    b->src[i] = s;
  }
  b->nsrc = n;
  return WQL_OK;
}

extern "C" int wql_block_set_coordinates(wql_block* b, const double* X, int32_t on_device) {
  if (!b || !X) return fail(WQL_EINVAL, "null argument");
  const long long n = (long long)b->nx * b->ny * b->nz;
  if (!b->X) {
    int rc = dalloc(b, (void**)&b->X, 3 * (size_t)b->P * sizeof(double));
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int c = 0; c < 3; ++c)
    HIP_TRY(hipMemcpy(b->X + c * b->P, X + c * n, (size_t)n * sizeof(double),
                      on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  return WQL_OK;
}

extern "C" int wql_block_set_receivers(wql_block* b, int32_t n, const int32_t* ijk, int32_t nslots) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (n < 0 || nslots < 0) return fail(WQL_EINVAL, "negative receiver count");
  if (b->rec_idx) { (void)hipFree(b->rec_idx); b->bytes -= b->nrec * (long long)sizeof(long long); b->rec_idx = nullptr; }
  if (b->seis) { (void)hipFree(b->seis); b->bytes -= 9LL * b->nrec * b->nslots * (long long)sizeof(double); b->seis = nullptr; }
  b->nrec = 0;
  b->nslots = 0;
  b->slot = 0;
  if (n == 0) return WQL_OK;
  std::vector<long long> idx(n);
  for (int r = 0; r < n; ++r) {
    const int x = ijk[3 * r] - 1, y = ijk[3 * r + 1] - 1, z = ijk[3 * r + 2] - 1;
    if (x < 0 || x >= b->nx || y < 0 || y >= b->ny || z < 0 || z >= b->nz)
      return fail(WQL_EINVAL, "receiver %d outside the block", r);
    idx[r] = x + (long long)b->nx * (y + (long long)b->ny * z);
  }
  int rc = dalloc(b, (void**)&b->rec_idx, n * sizeof(long long));
  if (rc) return rc;
  rc = dalloc(b, (void**)&b->seis, (size_t)9 * n * (nslots > 0 ? nslots : 1) * sizeof(double));
  if (rc) return rc;
  HIP_TRY(hipMemcpy(b->rec_idx, idx.data(), n * sizeof(long long), hipMemcpyHostToDevice));
  b->nrec = n;
  b->nslots = nslots;
  return WQL_OK;
}

extern "C" int wql_block_get_seismograms(wql_block* b, int32_t nslots, double* out) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (nslots > b->nslots) return fail(WQL_EINVAL, "only %d slots allocated", b->nslots);
  if (nslots <= 0 || b->nrec == 0) return WQL_OK;
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(out, b->seis, (size_t)9 * b->nrec * nslots * sizeof(double),
                    hipMemcpyDeviceToHost));
  return WQL_OK;
}

// ---------------------------------------------------------------------------
static double source_time(int type, double t, double t0, double w) {
  const double pi = 3.141592653589793;
  switch (type) {
    case WQL_SRC_GAUSSIAN:  // moment_tensor.f90:829-833
      return (w / (sqrt(2.0 * pi))) * exp((-(w * w) * ((t - t0) * (t - t0))) / 2.0);
    case WQL_SRC_COSBELL:   // moment_tensor.f90:843-849
      return (t < t0) ? (1.0 - cos(2.0 * pi * t / t0)) / t0 : 0.0;
    case WQL_SRC_RICKER: {  // Ricker wavelet (no reference counterpart): w = peak frequency
      const double a = pi * w * (t - t0);
      return (1.0 - 2.0 * a * a) * exp(-a * a);
    }
  }
  return 0.0;
}

template <typename T, typename S = double>
static StepArgs<T, S> make_args(wql_block* b, double t) {
  const wql_block_desc& d = b->d;
  StepArgs<T, S> a;
  memset(&a, 0, sizeof a);
  a.nx = b->nx; a.ny = b->ny; a.nz = b->nz;
  a.P = b->P;
  a.nmech = d.nmech;
  a.h[0] = d.hq; a.h[1] = d.hr; a.h[2] = d.hs;
  for (int k = 0; k < 3; ++k) { a.ih[k] = 1.0 / a.h[k]; a.pen[k] = d.tau0 / a.h[k]; }
  a.Fin = (const S*)b->F[b->cur];
  a.Fout = (S*)b->F[b->cur ^ 1];
  a.Fo_in = (const T*)b->Fo[b->cur];
  a.Fo_out = (T*)b->Fo[b->cur ^ 1];
  a.DF = (S*)b->DF;
  a.Fc = (S*)b->Fc;
  a.G = (const T*)b->G;
  a.M = (const T*)b->M;
  a.eta = (T*)b->eta;
  a.Deta = (T*)b->Deta;
  a.Qs = (const T*)b->Qs;
  a.Qp = (const T*)b->Qp;
  for (int i = 0; i < d.nmech; ++i) { a.tau_inv[i] = 1.0 / d.tau[i]; a.wgt[i] = d.weight[i]; }
  for (int f = 0; f < 6; ++f) {
    a.pml[f] = d.pml[f];
    a.npml[f] = d.pml[f] ? d.npml[f] : 0;
    a.pQ[f] = (T*)b->pQ[f];
    a.pDQ[f] = (T*)b->pDQ[f];
    a.bc[f] = d.bc[f] > 0 ? d.bc[f] : 0;
    a.face[f] = (const T*)b->face[f];
    a.iface[f] = d.bc[f] == 0 ? b->iface_sat[f] : nullptr;
  }
  a.sat_form = d.sat_form;
  a.X = b->X;
  a.nsrc = b->nsrc;
  for (int s = 0; s < b->nsrc; ++s) {
    const wql_source& Sr = b->src[s];
    DevSource& D = a.src[s];
    for (int k = 0; k < 3; ++k) {
      D.near[k] = Sr.near[k] - 1;
      D.off[k] = Sr.location[k] - Sr.near_phys[k];
      D.loc[k] = Sr.location[k];
    }
    D.order = Sr.order;
    D.alpha = Sr.alpha;
    const double A = source_time(Sr.type, t, Sr.t_init, Sr.duration);
    const double m[6] = {Sr.mXX, Sr.mYY, Sr.mZZ, Sr.mXY, Sr.mXZ, Sr.mYZ};
    for (int c = 0; c < 6; ++c) D.am[c] = A * m[c];
  }
  // strict interior (JU_xJU_yJU_z6.f90:16158-16173), 0-based inclusive
  auto lo = [&](int f) { return d.pml[f] ? (d.npml[f] > 6 ? d.npml[f] : 6) : 6; };
  a.ix0 = lo(0); a.ix1 = b->nx - 1 - lo(1);
  a.iy0 = lo(2); a.iy1 = b->ny - 1 - lo(3);
  a.iz0 = lo(4); a.iz1 = b->nz - 1 - lo(5);
  a.band_rows = (b->tune_band > 0 && b->tune_band < b->ny) ? b->tune_band : b->ny;
  a.segs = (int)(((long long)a.band_rows * b->nx + 255) / 256);
  a.xcd_remap = b->tune_xcd;
  a.nontemporal = b->tune_nt;
  a.W = (T*)b->W;
  a.split = b->tune_split;
  a.zm = b->tune_zm;
  a.zchunk = b->tune_zchunk;
  a.node0 = 0;
  a.node1 = b->P;
  a.mem_npt = b->tune_mem_npt;
  a.mem_nt = b->tune_mem_nt;
  a.mem_xcd = b->tune_mem_xcd;
  // a zero-weight mechanism whose state is all zero stays zero:
  // Deta = A*0 + (0 - 0)/tau = 0, eta = 0 + Bdt*0 = 0
  int zero_weight = 0;
  for (int i = 0; i < d.nmech; ++i)
    if (d.weight[i] == 0.0) zero_weight |= 1 << i;
  a.mech_skip = zero_weight & b->mech_zero;
  return a;
}

template <typename T, typename S>
static hipError_t arena_probe_launch(wql_block* b) {
  StepArgs<T, S> a = make_args<T, S>(b, 0.0);
  // A != 0: the variant of stages 2..5, which also reads the old rates
  // (all zero here, so every plane stays zero)
  a.A = 1.0;
  a.Bdt = 0.0;
  return launch_stage_part_z<T, S>(a, b->d.mesh, StagePart::kMemory, 0, b->nz, b->stream);
}

static int arena_probe_ms(wql_block* b, void* arena, size_t bytes, float* ms) {
  // the zero-weight mechanisms are skipped after wql_block_init_state;
  // probe with the same streams
  const int mz = b->mech_zero;
  b->mech_zero = ~0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMemsetAsync(arena, 0, bytes, b->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    if (rep == 1) e = hipEventRecord(e0, b->stream);
    if (e != hipSuccess) break;
    switch (b->d.precision) {
      case WQL_FP64: e = arena_probe_launch<double, double>(b); break;
      case WQL_FP32_MIXED: e = arena_probe_launch<float, double>(b); break;
      default: e = arena_probe_launch<float, float>(b); break;
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  b->mech_zero = mz;
  if (e != hipSuccess) return fail(WQL_EHIP, "arena probe: %s", hipGetErrorString(e));
  return WQL_OK;
}

template <typename T>
static int scale_t(wql_block* b, double A) {
  const long long P = b->P;
  HIP_TRY(launch_scale<double>((double*)b->DF, 9 * P, A, b->stream));
  if (b->d.nmech) HIP_TRY(launch_scale<T>((T*)b->Deta, 6LL * b->d.nmech * P, A, b->stream));
  for (int f = 0; f < 6; ++f)
    if (b->pDQ[f]) HIP_TRY(launch_scale<T>((T*)b->pDQ[f], 9 * b->Ps[f], A, b->stream));
  return WQL_OK;
}

template <typename T>
static int update_t(wql_block* b, double inc) {
  const long long P = b->P;
  b->fo_valid = false;
  HIP_TRY(launch_update<double>((double*)b->F[b->cur], (const double*)b->DF, 9 * P, inc, b->stream));
  if (b->d.nmech)
    HIP_TRY(launch_update<T>((T*)b->eta, (const T*)b->Deta, 6LL * b->d.nmech * P, inc, b->stream));
  for (int f = 0; f < 6; ++f)
    if (b->pQ[f]) HIP_TRY(launch_update<T>((T*)b->pQ[f], (const T*)b->pDQ[f], 9 * b->Ps[f], inc, b->stream));
  return WQL_OK;
}

static int no_kahan(wql_block* b, const char* what) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (b->d.precision == WQL_FP32_KAHAN)
    return fail(WQL_ESTATE, "%s needs an FP64 RK state; the FP32-Kahan state runs through "
                "wql_stage_fused / wql_time_step", what);
  return WQL_OK;
}

extern "C" int wql_rk_scale(wql_block* b, double A) {
  if (int rc = no_kahan(b, "wql_rk_scale")) return rc;
  return b->d.precision == WQL_FP64 ? scale_t<double>(b, A) : scale_t<float>(b, A);
}

extern "C" int wql_rk_update(wql_block* b, double inc) {
  if (int rc = no_kahan(b, "wql_rk_update")) return rc;
  return b->d.precision == WQL_FP64 ? update_t<double>(b, inc) : update_t<float>(b, inc);
}

extern "C" int wql_scale_rate_array(double* rate, int64_t n, double c) {
  if (n < 0 || (n > 0 && !rate)) return fail(WQL_EINVAL, "bad array");
  HIP_TRY(launch_scale<double>(rate, n, c, 0));
  HIP_TRY(hipStreamSynchronize(0));
  return WQL_OK;
}

extern "C" int wql_update_state_array(double* state, const double* rate, int64_t n, double inc) {
  if (n < 0 || (n > 0 && (!state || !rate))) return fail(WQL_EINVAL, "bad array");
  HIP_TRY(launch_update<double>(state, rate, n, inc, 0));
  HIP_TRY(hipStreamSynchronize(0));
  return WQL_OK;
}

// ---------------------------------------------------------------------------
// device mirrors of single host arrays: the present table of the reference's
// OpenACC backend (acc enter data / acc_is_present / update device / update
// self / exit data, rk_vector_openacc.f90, tests/rk_vector_test_data_openacc.f90)
// ---------------------------------------------------------------------------
struct Mirror {
  uintptr_t host;
  size_t bytes;
  void* dev;
};
static std::mutex g_mirror_mu;
static std::vector<Mirror> g_mirrors;

// the mirror holding [host, host + bytes), or null (caller holds the lock)
static Mirror* find_mirror(const void* host, size_t bytes) {
  const uintptr_t h = (uintptr_t)host;
  for (Mirror& m : g_mirrors)
    if (h >= m.host && h + bytes <= m.host + m.bytes) return &m;
  return nullptr;
}

extern "C" int wql_array_enter(const void* host, int64_t bytes) {
  if (!host || bytes <= 0) return fail(WQL_EINVAL, "bad host array");
  std::lock_guard<std::mutex> lock(g_mirror_mu);
  if (find_mirror(host, (size_t)bytes)) return WQL_OK;  // already present
  for (const Mirror& m : g_mirrors)
    if ((uintptr_t)host < m.host + m.bytes && m.host < (uintptr_t)host + (size_t)bytes)
      return fail(WQL_EINVAL, "host range overlaps a present array only in part");
  Mirror m{(uintptr_t)host, (size_t)bytes, nullptr};
  hipError_t e = hipMalloc(&m.dev, m.bytes);
  if (e != hipSuccess) return fail(WQL_ENOMEM, "hipMalloc(%zu): %s", m.bytes, hipGetErrorString(e));
  e = hipMemcpy(m.dev, host, m.bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(m.dev);
    return fail(WQL_EHIP, "copy in: %s", hipGetErrorString(e));
  }
  g_mirrors.push_back(m);
  return WQL_OK;
}

extern "C" int wql_array_exit(const void* host) {
  std::lock_guard<std::mutex> lock(g_mirror_mu);
  for (size_t k = 0; k < g_mirrors.size(); ++k)
    if (g_mirrors[k].host == (uintptr_t)host) {
      HIP_TRY(hipDeviceSynchronize());
      (void)hipFree(g_mirrors[k].dev);
      g_mirrors.erase(g_mirrors.begin() + (long)k);
      return WQL_OK;
    }
  return fail(WQL_EINVAL, "array at %p is not present", host);
}

extern "C" int wql_array_present(const void* host, int64_t bytes) {
  if (!host || bytes < 0) return 0;
  std::lock_guard<std::mutex> lock(g_mirror_mu);
  return find_mirror(host, (size_t)bytes) ? 1 : 0;
}

static int mirror_copy(void* host, int64_t bytes, bool to_device) {
  if (!host || bytes < 0) return fail(WQL_EINVAL, "bad host array");
  std::lock_guard<std::mutex> lock(g_mirror_mu);
  Mirror* m = find_mirror(host, (size_t)bytes);
  if (!m) return fail(WQL_ESTATE, "array at %p (%lld bytes) is not present on the device", host,
                      (long long)bytes);
  char* d = (char*)m->dev + ((uintptr_t)host - m->host);
  if (to_device) HIP_TRY(hipMemcpy(d, host, (size_t)bytes, hipMemcpyHostToDevice));
  else HIP_TRY(hipMemcpy(host, d, (size_t)bytes, hipMemcpyDeviceToHost));
  return WQL_OK;
}

extern "C" int wql_array_update_device(const void* host, int64_t bytes) {
  return mirror_copy((void*)host, bytes, true);
}

extern "C" int wql_array_update_self(void* host, int64_t bytes) {
  return mirror_copy(host, bytes, false);
}

// the device address of a present host range
static int mirror_of(const void* host, size_t bytes, double** dev) {
  std::lock_guard<std::mutex> lock(g_mirror_mu);
  Mirror* m = find_mirror(host, bytes);
  if (!m) return fail(WQL_ESTATE, "RK array at %p (%zu bytes) is not present on the device",
                      host, bytes);
  *dev = (double*)((char*)m->dev + ((uintptr_t)host - m->host));
  return WQL_OK;
}

extern "C" int wql_scale_rate_present(double* rate, int64_t n, double coefficient) {
  if (!rate || n < 0) return fail(WQL_EINVAL, "bad array");
  double* d = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  if (int rc = mirror_of(rate, bytes, &d)) return rc;
  HIP_TRY(hipMemcpy(d, rate, bytes, hipMemcpyHostToDevice));   // update device(rate)
  HIP_TRY(launch_scale<double>(d, n, coefficient, 0));
  HIP_TRY(hipMemcpy(rate, d, bytes, hipMemcpyDeviceToHost));   // update self(rate)
  return WQL_OK;
}

extern "C" int wql_update_state_present(double* state, const double* rate, int64_t n,
                                        double increment) {
  if (!state || !rate || n < 0) return fail(WQL_EINVAL, "bad array");
  double *ds = nullptr, *dr = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  if (int rc = mirror_of(state, bytes, &ds)) return rc;
  if (int rc = mirror_of(rate, bytes, &dr)) return rc;
  HIP_TRY(hipMemcpy(ds, state, bytes, hipMemcpyHostToDevice));  // update device(state)
  HIP_TRY(hipMemcpy(dr, rate, bytes, hipMemcpyHostToDevice));   // update device(rate)
  HIP_TRY(launch_update<double>(ds, dr, n, increment, 0));
  HIP_TRY(hipMemcpy(state, ds, bytes, hipMemcpyDeviceToHost));  // update self(state)
  return WQL_OK;
}

static int need_fp64(wql_block* b, const char* what) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (b->d.precision != WQL_FP64)
    return fail(WQL_ESTATE, "%s is the FP64 reference-order path; use wql_stage_fused for FP32 modes",
                what);
  return WQL_OK;
}

extern "C" int wql_rhs(wql_block* b) {
  int rc = need_fp64(b, "wql_rhs");
  if (rc) return rc;
  HIP_TRY(launch_rhs(make_args<double>(b, 0.0), b->d.mesh, RhsPart::kAll, b->stream));
  return WQL_OK;
}

extern "C" int wql_rhs_interior(wql_block* b, int32_t* handled) {
  if (handled) *handled = 0;
  int rc = need_fp64(b, "wql_rhs_interior");
  if (rc) return rc;
  HIP_TRY(launch_rhs(make_args<double>(b, 0.0), b->d.mesh, RhsPart::kInterior, b->stream));
  if (handled) *handled = 1;
  return WQL_OK;
}

extern "C" int wql_set_moment_tensor(wql_block* b, double t) {
  int rc = need_fp64(b, "wql_set_moment_tensor");
  if (rc) return rc;
  if (b->nsrc) HIP_TRY(launch_source(make_args<double>(b, t), b->d.mesh, b->stream));
  return WQL_OK;
}

extern "C" int wql_sat(wql_block* b) {
  int rc = need_fp64(b, "wql_sat");
  if (rc) return rc;
  HIP_TRY(launch_sat(make_args<double>(b, 0.0), b->stream));
  return WQL_OK;
}

static int record(wql_block* b) {
  if (b->nrec && b->slot < b->nslots) {
    if (b->d.precision == WQL_FP32_KAHAN)
      HIP_TRY(launch_sample<float>((const float*)b->F[b->cur], b->P, b->rec_idx, b->nrec,
                                   b->seis + (size_t)9 * b->nrec * b->slot, b->stream));
    else
      HIP_TRY(launch_sample<double>((const double*)b->F[b->cur], b->P, b->rec_idx, b->nrec,
                          b->seis + (size_t)9 * b->nrec * b->slot, b->stream));
    ++b->slot;
  }
  return WQL_OK;
}

// next free timing event of the block (grows the pool by 128 events at a time)
static int next_event(wql_block* b, hipEvent_t* out) {
  if (b->nev_used == b->nev_alloc) {
    const int n = b->nev_alloc + 128;
    hipEvent_t* ev = (hipEvent_t*)realloc(b->ev, sizeof(hipEvent_t) * n);
    if (!ev) return fail(WQL_ENOMEM, "host allocation failed");
    b->ev = ev;
    for (int i = b->nev_alloc; i < n; ++i) {
      HIP_TRY(hipEventCreate(&b->ev[i]));
      b->nev_alloc = i + 1;
    }
  }
  *out = b->ev[b->nev_used++];
  return WQL_OK;
}

static int join(wql_block* b) {
  if (b->pending) {
    HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_mem[b->nchunks - 1], 0));
    b->pending = 0;
  }
  return WQL_OK;
}

// z levels [z0, z1) of chunk k of K over nz levels
static void chunk_range(int nz, int K, int k, int* z0, int* z1) {
  *z0 = (int)((long long)nz * k / K);
  *z1 = (int)((long long)nz * (k + 1) / K);
}

// The chunk whose previous-stage memory kernel the stencils of chunk k must
// wait for: the chunk holding the highest z level they read.  The central
// stencil reads z1 + 2 (3-node halo); the SBP closure rows z = 0..5 read
// taps up to z = 8 (row 6 of the closure has 9 taps), so a chunk with a
// bottom closure row reads up to max(z1 + 2, 8).  Levels below z0 belong to
// lower chunks, which stream2 has finished before (it runs in order).
static int chunk_wait_index(int nz, int K, int k) {
  int z0, z1;
  chunk_range(nz, K, k, &z0, &z1);
  int zmax = z1 + 2;
  if (z0 < 6 && zmax < 8) zmax = 8;
  if (zmax > nz - 1) zmax = nz - 1;
  int j = k;
  for (;;) {
    int c0, c1;
    chunk_range(nz, K, j, &c0, &c1);
    if (zmax < c1 || j == K - 1) break;
    ++j;
  }
  return j;
}

// One stage as a z-chunk pipeline: the stencil kernels (latency bound) of
// chunk k on the block's stream overlap the memory-variable kernel (HBM
// bound) of chunk k - 1 on stream2.  Hazards: the stencil of chunk k reads F
// of levels z0-3 .. max(z1+2, 8 for a chunk with bottom closure rows),
// written by the previous stage's memory kernels of the chunks holding them
// (it waits for the event of the chunk holding the highest of these levels,
// chunk_wait_index; stream2 runs in order, so the lower chunks are done); it
// writes W and the PML fields of chunk k only, which the previous stage's
// memory kernel of chunk k has finished reading (same wait).  The memory
// kernel of chunk k writes F_out, DF, eta of chunk k after the stencil of
// chunk k (ev_sten[k]); F_out is the buffer this stage's stencils do not read.
// The stencils of all chunks run in order on the block's stream, so those of
// the previous stage (whose top closure rows read down to z = nz - 9) have
// finished before this stage's stencils write the velocity part of F_out.
template <typename T, typename S>
static int stage_pipelined(wql_block* b, const StepArgs<T, S>& a) {
  const int K = b->nchunks, nz = b->nz;
  for (int k = 0; k < K; ++k) {
    int z0, z1;
    chunk_range(nz, K, k, &z0, &z1);
    const int kw = chunk_wait_index(nz, K, k);
    {
      // the waited-for chunk ends above every level this chunk's stencils read
      int w0, w1;
      chunk_range(nz, K, kw, &w0, &w1);
      const int need = std::min(nz - 1, std::max(z1 + 2, z0 < 6 ? 8 : 0));
      if (need >= w1) return fail(WQL_ESTATE, "z-chunk wait rule: chunk %d reads level %d", k, need);
    }
    HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_mem[kw], 0));
    HIP_TRY(launch_stage_part_z<T>(a, b->d.mesh, StagePart::kInterior, z0, z1, b->stream));
    HIP_TRY(launch_stage_part_z<T>(a, b->d.mesh, StagePart::kShell, z0, z1, b->stream));
    HIP_TRY(hipEventRecord(b->ev_sten[k], b->stream));
    HIP_TRY(hipStreamWaitEvent(b->stream2, b->ev_sten[k], 0));
    HIP_TRY(launch_stage_part_z<T>(a, b->d.mesh, StagePart::kMemory, z0, z1, b->stream2));
    HIP_TRY(hipEventRecord(b->ev_mem[k], b->stream2));
  }
  b->pending = 1;
  return WQL_OK;
}

// FP32-mixed: the float copy of F[cur] that every kernel differencing F reads
// (the fused stage's stencils, the adjoint's material gradient), rebuilt here
// after F changed outside the fused stage; every fused stage writes Fo_out
// beside F_out, so the copy of the new state is current afterwards
static int ensure_operand_copy(wql_block* b) {
  if (!b->Fo[0] || b->fo_valid) return WQL_OK;
  if (int rc = join(b)) return rc;
  HIP_TRY(launch_round_float((float*)b->Fo[b->cur], (const double*)b->F[b->cur], 9 * b->P,
                             b->stream));
  b->fo_valid = true;
  return WQL_OK;
}

template <typename T, typename S = double>
static int stage_fused_t(wql_block* b, double A, double Bdt, double t, bool keep_pending = false,
                         bool dead_rates = false) {
  StepArgs<T, S> a = make_args<T, S>(b, t);
  a.A = A; a.Bdt = Bdt;
  a.dead_rates = dead_rates ? 1 : 0;
  if (int rc = ensure_operand_copy(b)) return rc;
  if (!b->timing) {
    if (b->nchunks > 1 && stage_has_memory_part(a)) {
      if (int rc = stage_pipelined<T, S>(b, a)) return rc;
      return keep_pending ? WQL_OK : join(b);
    }
    if (int rc = join(b)) return rc;
    HIP_TRY(launch_stage_fused<T>(a, b->d.mesh, b->stream));
    return WQL_OK;
  }
  if (int rc = join(b)) return rc;
  hipEvent_t e[4];
  for (int k = 0; k < 4; ++k) {
    int rc = next_event(b, &e[k]);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(e[0], b->stream));
  const StagePart parts[3] = {StagePart::kInterior, StagePart::kShell, StagePart::kMemory};
  for (int k = 0; k < 3; ++k) {
    HIP_TRY(launch_stage_part<T>(a, b->d.mesh, parts[k], b->stream));
    HIP_TRY(hipEventRecord(e[k + 1], b->stream));
  }
  return WQL_OK;
}

// one fused stage; keep_pending leaves the pipelined memory kernels of the
// stage unjoined (the next stage of the same block orders itself behind them)
static int stage_fused(wql_block* b, double A, double Bdt, double t, bool keep_pending,
                       bool dead_rates = false) {
  int rc = b->d.precision == WQL_FP64 ? stage_fused_t<double>(b, A, Bdt, t, keep_pending, dead_rates)
           : b->d.precision == WQL_FP32_MIXED
               ? stage_fused_t<float>(b, A, Bdt, t, keep_pending, dead_rates)
               : stage_fused_t<float, float>(b, A, Bdt, t, keep_pending, dead_rates);
  if (rc) return rc;
  b->cur ^= 1;
  return WQL_OK;
}

extern "C" int wql_stage_fused(wql_block* b, double A, double Bdt, double t) {
  if (!b) return fail(WQL_EINVAL, "null block");
  return stage_fused(b, A, Bdt, t, false);
}

extern "C" int wql_block_set_timing(wql_block* b, int32_t enable) {
  if (!b) return fail(WQL_EINVAL, "null block");
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->timing = enable != 0;
  b->nev_used = 0;
  return WQL_OK;
}

extern "C" int wql_block_kernel_times(wql_block* b, double* ms, int32_t* stages) {
  if (!b || !ms || !stages) return fail(WQL_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(b->stream));
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k + 3 < b->nev_used; k += 4) {
    for (int j = 0; j < 3; ++j) {
      float m = 0.f;
      HIP_TRY(hipEventElapsedTime(&m, b->ev[k + j], b->ev[k + j + 1]));
      acc[j] += m;
    }
  }
  for (int j = 0; j < 3; ++j) ms[j] = acc[j];
  *stages = b->nev_used / 4;
  b->nev_used = 0;
  return WQL_OK;
}

extern "C" int wql_time_step(wql_block* b, double dt, int32_t nsteps, double* t) {
  if (!b || !t) return fail(WQL_EINVAL, "null argument");
  for (int n = 0; n < nsteps; ++n) {
    const double t0 = *t;
    for (int s = 0; s < 5; ++s) {
      if (s == 0 && b->nrec && b->slot < b->nslots) {
        int rc = join(b);
        if (!rc) rc = record(b);  // receivers sample F at t0 (time_step.f90:137)
        if (rc) return rc;
      }
      // the last stage's rates are dead when another step of this call
      // follows: its first stage has A = 0 and reads none of them
      const bool dead = s == 4 && n + 1 < nsteps && kRkA[0] == 0.0;
      int rc = stage_fused(b, kRkA[s], kRkB[s] * dt, *t, true, dead);
      if (rc) return rc;
      *t = t0 + kRkC[s + 1] * dt;
    }
  }
  return join(b);
}

// ---------------------------------------------------------------------------
// two-block interface
// ---------------------------------------------------------------------------
static int iface_ok(wql_iface* I) {
  if (!I) return fail(WQL_EINVAL, "null interface");
  if (!I->L || !I->R) return fail(WQL_ESTATE, "a block of the interface was destroyed");
  return WQL_OK;
}

static double boxcar(double x, double w) {
  const double tol = 100.0 * 2.220446049250313e-16;  // Interface_Condition.f90:2438-2457
  if (-w + tol < x && x < w - tol) return 1.0;
  if (fabs(-w - x) <= tol || fabs(x - w) <= tol) return 0.5;
  return 0.0;
}

// copy NF doubles (host or device source) to host
static int to_host(std::vector<double>& dst, const double* src, long long n, bool on_device) {
  dst.resize((size_t)n);
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
  HIP_TRY(hipMemcpy(dst.data(), src, (size_t)n * sizeof(double), kind));
  return WQL_OK;
}

// the FP64 material (lambda, mu, rho) of the nodes of q face f (0: x = 0,
// 1: x = nx - 1) from the descriptor's array (host or device), [c][y + ny z]
static int face_material64(wql_block* b, int f, std::vector<double>& out) {
  const wql_block_desc& d = b->d;
  if (!d.M) return fail(WQL_EINVAL, "material array M is null");
  const long long NF = (long long)b->ny * b->nz;
  out.assign((size_t)(3 * NF), 0.0);
  const View v = make_view(d.M, d.body);
  const long long x1 = f == 1 ? b->nx : 1;
  for (int c = 0; c < 3; ++c) {
    const long long off0 = view_off(v, x1, 1, 1, c);
    if (d.arrays_on_device) {
      double* tmpd = nullptr;
      HIP_TRY(hipMalloc((void**)&tmpd, NF * sizeof(double)));
      hipError_t e = launch_gather<double>(tmpd, d.M + off0, 1, b->ny, b->nz, 1, v.sx, v.sy,
                                           v.sz, 0, 0, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e == hipSuccess)
        e = hipMemcpy(out.data() + c * NF, tmpd, NF * sizeof(double), hipMemcpyDeviceToHost);
      (void)hipFree(tmpd);
      if (e != hipSuccess) return fail(WQL_EHIP, "face material gather: %s", hipGetErrorString(e));
    } else {
      for (long long z = 0; z < b->nz; ++z)
        for (long long y = 0; y < b->ny; ++y)
          out[c * NF + y + b->ny * z] = d.M[off0 + y * v.sy + z * v.sz];
    }
  }
  return WQL_OK;
}

// FP64 copies of the coupled faces' frames and material in the interface
// buffer (in every precision the interface arithmetic is FP64; the blocks'
// FP32 planes are not its operands)
static int iface_fill_faces(wql_iface* I) {
  wql_block* L = I->L;
  wql_block* R = I->R;
  const long long NF = I->NF;
  if (!L->face64[1] || !R->face64[0])
    return fail(WQL_ESTATE, "the coupled faces' frames have not been uploaded (WQL_FACES)");
  const size_t nb = (size_t)NF * sizeof(double);
  HIP_TRY(hipMemcpy(I->frL, L->face64[1], 13 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->frR, R->face64[0], 13 * nb, hipMemcpyHostToDevice));
  std::vector<double> mL, mR;
  if (int rc = face_material64(L, 1, mL)) return rc;
  if (int rc = face_material64(R, 0, mR)) return rc;
  HIP_TRY(hipMemcpy(I->matL, mL.data(), 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->matR, mR.data(), 3 * nb, hipMemcpyHostToDevice));
  return WQL_OK;
}

static bool is_rate_state(int coupling) {
  return coupling == WQL_IFACE_RATE_STATE_TPV101 || coupling == WQL_IFACE_RATE_STATE_TPV102 ||
         coupling == WQL_IFACE_RATE_STATE_RW;
}

// smooth_boxcar (Interface_Condition.f90:2460-2476)
static double smooth_boxcar(double x, double w, double wr) {
  if (-w <= x && x <= w) return 1.0;
  if (fabs(x) > w && fabs(x) < w + wr)
    return 0.5 * (1.0 + tanh(wr / (fabs(x) - w - wr) + wr / (fabs(x) - w)));
  return 0.0;
}

// Slip-weakening tables of the interface: prestress in each side's (n, m, l)
// frame and fs, fd, Dc per node (see wql3d.h).
static int iface_tables(wql_iface* I, wql_block* L, wql_block* R, const wql_iface_desc* d) {
  const long long NF = I->NF;
  if (d->coupling == WQL_IFACE_LOCKED) return WQL_OK;
  const bool dev = d->arrays_on_device != 0;
  std::vector<double> T0L(3 * NF), T0R(3 * NF), fs(NF), fd(NF), Dc(NF);
  if (d->coupling == WQL_IFACE_LINEAR_FRICTION) {
    // no friction tables; the prestress the fault output adds, if given
    if (d->T0L && d->T0R) {
      if (int rc = to_host(T0L, d->T0L, 3 * NF, dev)) return rc;
      if (int rc = to_host(T0R, d->T0R, 3 * NF, dev)) return rc;
    } else {
      std::fill(T0L.begin(), T0L.end(), 0.0);
      std::fill(T0R.begin(), T0R.end(), 0.0);
    }
    const size_t nb = (size_t)NF * sizeof(double);
    HIP_TRY(hipMemcpy(I->T0L, T0L.data(), 3 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->T0R, T0R.data(), 3 * nb, hipMemcpyHostToDevice));
    return WQL_OK;
  }
  if (d->coupling == WQL_IFACE_RATE_STATE_RW) {
    // problem 'rate-weakening' (Interface_Condition.f90:410-416, 438-444,
    // 2387-2426): the traction of initial_stress_tensor('rate-weakening')
    // (initial_stress_condition.f90:184-191, with its pi = 3.14159265359) in
    // each side's frame, a(y, z) and V_w(y, z) with smooth_boxcar(20, 2.5),
    // and the pulse shape F(r) = exp(-(r/r0)^2/2), r0 = 3*1.5*0.3, around
    // (y, z) = (8, 0) (the stage adds 2 (-0.2429 T0_n) F(r) to T0_l for t > 0)
    std::vector<double> y, z, a(NF), vw(NF), nuc(NF), W0(NF, 0.0);
    if (int rc = to_host(y, d->y, NF, dev)) return rc;
    if (int rc = to_host(z, d->z, NF, dev)) return rc;
    if (d->W0)
      if (int rc = to_host(W0, d->W0, NF, dev)) return rc;
    const double pi_s = 3.14159265359;
    const double psi = 50.0 / 180.0 * pi_s;
    double s0[6];
    s0[0] = -126.0;
    s0[4] = 42.0;
    s0[2] = (1.0 - 2.0 * s0[4] / (s0[0] * tan(2.0 * psi))) * s0[0];
    s0[1] = 0.5 * (s0[0] + s0[2]);
    s0[3] = 0.0;
    s0[5] = 0.0;
    const double* frs[2] = {L->face64[1], R->face64[0]};
    std::vector<double>* outs[2] = {&T0L, &T0R};
    const double r0 = 3.0 * 1.5 * 0.3;
    for (long long k = 0; k < NF; ++k) {
      for (int side = 0; side < 2; ++side) {
        const double* fr = frs[side];
        double l[3], m[3], n[3];
        for (int c = 0; c < 3; ++c) {
          l[c] = fr[c * NF + k];
          m[c] = fr[(3 + c) * NF + k];
          n[c] = fr[(6 + c) * NF + k];
        }
        const double Tx = n[0] * s0[0] + n[1] * s0[3] + n[2] * s0[4];
        const double Ty = n[0] * s0[3] + n[1] * s0[1] + n[2] * s0[5];
        const double Tz = n[0] * s0[4] + n[1] * s0[5] + n[2] * s0[2];
        std::vector<double>& t = *outs[side];
        t[k] = n[0] * Tx + n[1] * Ty + n[2] * Tz;
        t[NF + k] = m[0] * Tx + m[1] * Ty + m[2] * Tz;
        t[2 * NF + k] = l[0] * Tx + l[1] * Ty + l[2] * Tz;
      }
      const double box = smooth_boxcar(y[k], 20.0, 2.5) * smooth_boxcar(z[k], 20.0, 2.5);
      a[k] = 0.016 + 0.016 * (1.0 - box);
      vw[k] = 0.17 + 5.0 * (1.0 - box);
      const double r = sqrt((y[k] - 8.0) * (y[k] - 8.0) + (z[k] + 0.0) * (z[k] + 0.0));
      nuc[k] = exp(-((r / r0) * (r / r0)) / 2.0);
    }
    const size_t nb = (size_t)NF * sizeof(double);
    HIP_TRY(hipMemcpy(I->T0L, T0L.data(), 3 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->T0R, T0R.data(), 3 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->rs_a, a.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->rs_nuc, nuc.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->rs_vw, vw.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->W, W0.data(), nb, hipMemcpyHostToDevice));
    return WQL_OK;
  }
  if (is_rate_state(d->coupling)) {
    // TPV101 / TPV102 (Interface_Condition.f90:386-400, 2324-2378): the
    // traction of the initial stress (-120, 0, -120, 0, 75, 0)
    // (initial_stress_condition.f90:114-120) in each side's frame, the
    // parameter a(y, z) and the nucleation amplitude 25 F(r) (the stage adds
    // 25 F(r) G(t) to T0_l); the initial state from the caller or 0
    // (iface.f90:38)
    std::vector<double> y, z, a(NF), nuc(NF), W0(NF, 0.0);
    if (int rc = to_host(y, d->y, NF, dev)) return rc;
    if (int rc = to_host(z, d->z, NF, dev)) return rc;
    if (d->W0)
      if (int rc = to_host(W0, d->W0, NF, dev)) return rc;
    static const double s0[6] = {-120.0, 0.0, -120.0, 0.0, 75.0, 0.0};
    const double* frs[2] = {L->face64[1], R->face64[0]};
    std::vector<double>* outs[2] = {&T0L, &T0R};
    for (long long k = 0; k < NF; ++k) {
      for (int side = 0; side < 2; ++side) {
        const double* fr = frs[side];
        double l[3], m[3], n[3];
        for (int c = 0; c < 3; ++c) {
          l[c] = fr[c * NF + k];
          m[c] = fr[(3 + c) * NF + k];
          n[c] = fr[(6 + c) * NF + k];
        }
        const double Tx = n[0] * s0[0] + n[1] * s0[3] + n[2] * s0[4];
        const double Ty = n[0] * s0[3] + n[1] * s0[1] + n[2] * s0[5];
        const double Tz = n[0] * s0[4] + n[1] * s0[5] + n[2] * s0[2];
        std::vector<double>& t = *outs[side];
        t[k] = n[0] * Tx + n[1] * Ty + n[2] * Tz;
        t[NF + k] = m[0] * Tx + m[1] * Ty + m[2] * Tz;
        t[2 * NF + k] = l[0] * Tx + l[1] * Ty + l[2] * Tz;
      }
      a[k] = 0.008 + 0.008 * (1.0 - smooth_boxcar(y[k], 15.0, 3.0) * smooth_boxcar(z[k], 15.0, 3.0));
      const double r = sqrt((y[k] - 7.5) * (y[k] - 7.5) + z[k] * z[k]);
      nuc[k] = r < 3.0 ? 25.0 * exp(r * r / (r * r - 9.0)) : 0.0;
    }
    const size_t nb = (size_t)NF * sizeof(double);
    HIP_TRY(hipMemcpy(I->T0L, T0L.data(), 3 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->T0R, T0R.data(), 3 * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->rs_a, a.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->rs_nuc, nuc.data(), nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(I->W, W0.data(), nb, hipMemcpyHostToDevice));
    return WQL_OK;
  }
  if (d->coupling == WQL_IFACE_SLIP_WEAKENING) {
    if (int rc = to_host(T0L, d->T0L, 3 * NF, dev)) return rc;
    if (int rc = to_host(T0R, d->T0R, 3 * NF, dev)) return rc;
    if (int rc = to_host(fs, d->fs, NF, dev)) return rc;
    if (int rc = to_host(fd, d->fd, NF, dev)) return rc;
    if (int rc = to_host(Dc, d->Dc, NF, dev)) return rc;
  } else {
    std::vector<double> y, z;
    if (int rc = to_host(y, d->y, NF, dev)) return rc;
    if (int rc = to_host(z, d->z, NF, dev)) return rc;
    if (d->coupling == WQL_IFACE_SLIP_WEAKENING_TPV5) {
      // prestress TPV5 (Interface_Condition.f90:1821-1830): T0 = (-120, 0, 70 + patches)
      for (long long k = 0; k < NF; ++k) {
        double l = 70.0;
        l = l + 8.0 * boxcar(y[k] - 7.5, 1.5) * boxcar(z[k] + 7.5, 1.5);
        l = l + 11.6 * boxcar(y[k] - 7.5, 1.5) * boxcar(z[k], 1.5);
        l = l - 8.0 * boxcar(y[k] - 7.5, 1.5) * boxcar(z[k] - 7.5, 1.5);
        for (std::vector<double>* t : {&T0L, &T0R}) {
          (*t)[k] = -120.0;
          (*t)[NF + k] = 0.0;
          (*t)[2 * NF + k] = l;
        }
        // Tau_strength TPV5 (:1283-1299)
        fs[k] = 0.677 + 1.0e10 * (1.0 - boxcar(y[k], 15.0) * boxcar(z[k], 15.0));
        fd[k] = 0.525;
        Dc[k] = 0.4;
      }
    } else {
      // prestress TPV28 (:2049-2079): traction of the initial stress on the
      // face in each side's frame, plus the nucleation stress on T0_l
      static const double kS0[6] = {-60.0, 0.0, 60.0, 0.0, 29.38, 0.0};
      const double* s0 = d->s0 ? d->s0 : kS0;
      const double pi = 3.141592653589793;
      std::vector<double> frL(L->face64[1], L->face64[1] + 9 * NF);
      std::vector<double> frR(R->face64[0], R->face64[0] + 9 * NF);
      for (long long k = 0; k < NF; ++k) {
        const double r = sqrt((y[k] - 7.5) * (y[k] - 7.5) + z[k] * z[k]);
        double tau_nuke;
        if (r <= 1.4) tau_nuke = 11.60;
        else if (r >= 1.4 && r <= 2.0) tau_nuke = 5.8 * (1.0 + cos(pi * (r - 1.4) / 0.6));
        else tau_nuke = 0.0;
        const std::vector<double>* frs[2] = {&frL, &frR};
        std::vector<double>* outs[2] = {&T0L, &T0R};
        for (int side = 0; side < 2; ++side) {
          const std::vector<double>& fr = *frs[side];
          double l[3], m[3], n[3];
          for (int c = 0; c < 3; ++c) {
            l[c] = fr[c * NF + k];
            m[c] = fr[(3 + c) * NF + k];
            n[c] = fr[(6 + c) * NF + k];
          }
          const double Tx = n[0] * s0[0] + n[1] * s0[3] + n[2] * s0[4];
          const double Ty = n[0] * s0[3] + n[1] * s0[1] + n[2] * s0[5];
          const double Tz = n[0] * s0[4] + n[1] * s0[5] + n[2] * s0[2];
          std::vector<double>& t = *outs[side];
          t[k] = n[0] * Tx + n[1] * Ty + n[2] * Tz;
          t[NF + k] = m[0] * Tx + m[1] * Ty + m[2] * Tz;
          t[2 * NF + k] = l[0] * Tx + l[1] * Ty + l[2] * Tz + tau_nuke;
        }
        // Tau_strength TPV28 (:1459-1475)
        fs[k] = 0.677 + 1.0e10 * (1.0 - boxcar(y[k], 15.0) * boxcar(z[k], 18.0));
        fd[k] = 0.373;
        Dc[k] = 0.4;
      }
    }
  }
  const size_t nb = (size_t)NF * sizeof(double);
  HIP_TRY(hipMemcpy(I->T0L, T0L.data(), 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->T0R, T0R.data(), 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->fs, fs.data(), nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->fd, fd.data(), nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(I->Dc, Dc.data(), nb, hipMemcpyHostToDevice));
  return WQL_OK;
}

extern "C" int wql_iface_create(wql_block* L, wql_block* R, const wql_iface_desc* d,
                                wql_iface** out) {
  if (!out) return fail(WQL_EINVAL, "null output pointer");
  *out = nullptr;
  if (!L || !R || !d || L == R) return fail(WQL_EINVAL, "need two distinct blocks and a descriptor");
  if (d->coupling < WQL_IFACE_LOCKED || d->coupling > WQL_IFACE_RATE_STATE_RW)
    return fail(WQL_EINVAL, "invalid coupling %d", d->coupling);
  if (d->coupling == WQL_IFACE_LINEAR_FRICTION && !std::isfinite(d->alpha))
    return fail(WQL_EINVAL, "linear friction needs a finite alpha (< 0: the reference's 1e14)");
  if ((d->coupling == WQL_IFACE_SLIP_WEAKENING_TPV5 || d->coupling == WQL_IFACE_SLIP_WEAKENING_TPV28 ||
       is_rate_state(d->coupling)) &&
      (!d->y || !d->z))
    return fail(WQL_EINVAL, "face coordinates y, z are null");
  if (d->coupling == WQL_IFACE_SLIP_WEAKENING &&
      (!d->T0L || !d->T0R || !d->fs || !d->fd || !d->Dc))
    return fail(WQL_EINVAL, "slip-weakening tables T0L, T0R, fs, fd, Dc are required");
  if (L->iface || R->iface) return fail(WQL_ESTATE, "a block is already coupled to an interface");
  if (L->d.precision != R->d.precision) return fail(WQL_EINVAL, "blocks differ in precision");
  if (L->d.bc[1] != 0 || R->d.bc[0] != 0)
    return fail(WQL_EINVAL, "the coupled faces (Rx of L, Lx of R) need bc 0 (got %d, %d)",
                L->d.bc[1], R->d.bc[0]);
  if (L->d.pml[1] || R->d.pml[0]) return fail(WQL_EINVAL, "the coupled faces cannot carry PML");
  if (L->ny != R->ny || L->nz != R->nz)
    return fail(WQL_EINVAL, "coupled faces differ: %dx%d vs %dx%d nodes", L->ny, L->nz, R->ny, R->nz);
  if (!L->face[1] || !R->face[0])
    return fail(WQL_EINVAL, "the coupled faces need n_l/n_m/n_n in the block descriptors");
  wql_iface* I = (wql_iface*)calloc(1, sizeof(wql_iface));
  if (!I) return fail(WQL_ENOMEM, "host allocation failed");
  I->L = L; I->R = R;
  I->ny = L->ny; I->nz = L->nz;
  I->NF = (long long)I->ny * I->nz;
  I->coupling = d->coupling;
  I->alpha = d->alpha < 0.0 ? 1.0e14 : d->alpha;
  const long long NF = I->NF;
  constexpr int kIfacePlanes = 36 + 5 + 26 + 6;
  hipError_t e = hipMalloc((void**)&I->buf, kIfacePlanes * NF * sizeof(double));
  if (e != hipSuccess) {
    free(I);
    return fail(WQL_ENOMEM, "hipMalloc(interface) failed: %s", hipGetErrorString(e));
  }
  I->S = I->buf; I->DS = I->S + 4 * NF; I->trup = I->DS + 4 * NF; I->T0L = I->trup + NF;
  I->T0R = I->T0L + 3 * NF; I->fs = I->T0R + 3 * NF; I->fd = I->fs + NF; I->Dc = I->fd + NF;
  I->satL = I->Dc + NF; I->satR = I->satL + 9 * NF;
  I->W = I->satR + 9 * NF; I->DW = I->W + NF; I->rs_a = I->DW + NF; I->rs_nuc = I->rs_a + NF;
  I->rs_vw = I->rs_nuc + NF;
  I->frL = I->rs_vw + NF; I->frR = I->frL + 13 * NF; I->matL = I->frR + 13 * NF;
  I->matR = I->matL + 3 * NF;
  auto cleanup = [&](int rc) {
    (void)hipFree(I->buf);
    if (I->ev_iface) (void)hipEventDestroy(I->ev_iface);
    if (I->ev_right) (void)hipEventDestroy(I->ev_right);
    free(I);
    return rc;
  };
  if ((e = launch_fill<double>(I->W, NF, 0.0, L->stream)) != hipSuccess ||
      (e = launch_fill<double>(I->rs_a, 3 * NF, 0.0, L->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(L->stream)) != hipSuccess)
    return cleanup(fail(WQL_EHIP, "interface init: %s", hipGetErrorString(e)));
  if (int rc = iface_fill_faces(I)) return cleanup(rc);
  if (int rc = iface_tables(I, L, R, d)) return cleanup(rc);
  // init_iface: S = 0, DS = 1e40, W as set above, DW = 1e40, trup = 1e9;
  // the SAT planes start at zero
  if ((e = launch_fill<double>(I->S, 4 * NF, 0.0, L->stream)) != hipSuccess ||
      (e = launch_fill<double>(I->DS, 4 * NF, 1.0e40, L->stream)) != hipSuccess ||
      (e = launch_fill<double>(I->DW, NF, 1.0e40, L->stream)) != hipSuccess ||
      (e = launch_fill<double>(I->trup, NF, 1.0e9, L->stream)) != hipSuccess ||
      (e = launch_fill<double>(I->satL, 18 * NF, 0.0, L->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(L->stream)) != hipSuccess)
    return cleanup(fail(WQL_EHIP, "interface init: %s", hipGetErrorString(e)));
  if ((e = hipEventCreateWithFlags(&I->ev_iface, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&I->ev_right, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventRecord(I->ev_right, R->stream)) != hipSuccess)
    return cleanup(fail(WQL_EHIP, "interface events: %s", hipGetErrorString(e)));
  L->iface = I; R->iface = I;
  L->iface_sat[1] = I->satL;
  R->iface_sat[0] = I->satR;
  *out = I;
  return WQL_OK;
}

extern "C" void wql_iface_destroy(wql_iface* I) {
  if (!I) return;
  if (I->L) {
    (void)hipStreamSynchronize(I->L->stream);
    I->L->iface = nullptr;
    I->L->iface_sat[1] = nullptr;
  }
  if (I->R) {
    (void)hipStreamSynchronize(I->R->stream);
    I->R->iface = nullptr;
    I->R->iface_sat[0] = nullptr;
  }
  if (I->buf) (void)hipFree(I->buf);
  if (I->frames) (void)hipFree(I->frames);
  if (I->ev_iface) (void)hipEventDestroy(I->ev_iface);
  if (I->ev_right) (void)hipEventDestroy(I->ev_right);
  free(I);
}

// one stage of the two-block domain: k_iface on L's stream (after R's
// previous stage), block L's fused stage behind it on L's stream, block R's on
// R's stream after k_iface
template <typename T, typename S = double>
static int iface_stage_t(wql_iface* I, double A, double Bdt, double t, int stage1,
                         bool dead_rates = false) {
  wql_block* L = I->L;
  wql_block* R = I->R;
  HIP_TRY(hipStreamWaitEvent(L->stream, I->ev_right, 0));
  IfaceArgs<T, S> a;
  memset(&a, 0, sizeof a);
  a.ny = I->ny; a.nz = I->nz;
  a.coupling = I->coupling;
  a.nxL = L->nx; a.nxR = R->nx;
  a.FL = (const S*)L->F[L->cur];
  a.FR = (const S*)R->F[R->cur];
  a.PL = L->P; a.PR = R->P;
  a.matL = I->matL; a.matR = I->matR;
  a.frL = I->frL; a.frR = I->frR;
  a.penL = L->d.tau0 / L->d.hq;
  a.penR = R->d.tau0 / R->d.hq;
  a.S = I->S; a.DS = I->DS; a.trup = I->trup;
  a.T0L = I->T0L; a.T0R = I->T0R; a.fs = I->fs; a.fd = I->fd; a.Dc = I->Dc;
  a.satL = I->satL; a.satR = I->satR;
  a.W = I->W; a.DW = I->DW; a.rs_a = I->rs_a; a.rs_nuc = I->rs_nuc; a.rs_vw = I->rs_vw;
  if (I->coupling == WQL_IFACE_RATE_STATE_RW) {
    // rate-weakening friction constants (Interface_Condition.f90:410-416)
    a.rs_L0 = 2.0 * 0.2572; a.rs_f0 = 0.7; a.rs_b = 0.02; a.rs_V0 = 1.0e-6;
  } else {
    // TPV101 / TPV102 friction constants (Interface_Condition.f90:386-400)
    a.rs_L0 = 0.02; a.rs_f0 = 0.6; a.rs_b = 0.012; a.rs_V0 = 1.0e-6;
  }
  a.alpha = I->alpha;
  a.A = A; a.Bdt = Bdt; a.t = t;
  a.stage1 = stage1 ? 1 : 0;
  a.frame = (stage1 && I->frames && I->slot < I->nslots)
                ? I->frames + (size_t)I->slot * kFramePlanes * I->NF : nullptr;
  HIP_TRY((launch_iface<T, S>(a, L->stream)));
  if (a.frame) ++I->slot;
  HIP_TRY(hipEventRecord(I->ev_iface, L->stream));
  int rc = stage_fused_t<T, S>(L, A, Bdt, t, false, dead_rates);
  if (rc) return rc;
  HIP_TRY(hipStreamWaitEvent(R->stream, I->ev_iface, 0));
  rc = stage_fused_t<T, S>(R, A, Bdt, t, false, dead_rates);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(I->ev_right, R->stream));
  // block L's stream then orders everything of the stage (callers time and
  // synchronise on it)
  HIP_TRY(hipStreamWaitEvent(L->stream, I->ev_right, 0));
  L->cur ^= 1;
  R->cur ^= 1;
  return WQL_OK;
}

static int iface_stage(wql_iface* I, double A, double Bdt, double t, int stage1, bool dead_rates) {
  switch (I->L->d.precision) {
    case WQL_FP64: return iface_stage_t<double>(I, A, Bdt, t, stage1, dead_rates);
    case WQL_FP32_MIXED: return iface_stage_t<float>(I, A, Bdt, t, stage1, dead_rates);
    default: return iface_stage_t<float, float>(I, A, Bdt, t, stage1, dead_rates);
  }
}

extern "C" int wql_iface_stage(wql_iface* I, double A, double Bdt, double t, int32_t stage1) {
  if (int rc = iface_ok(I)) return rc;
  return iface_stage(I, A, Bdt, t, stage1, false);
}

extern "C" int wql_iface_get_friction_state(wql_iface* I, double* W) {
  if (int rc = iface_ok(I)) return rc;
  if (!W) return fail(WQL_EINVAL, "null output");
  HIP_TRY(hipStreamSynchronize(I->L->stream));
  HIP_TRY(hipMemcpy(W, I->W, (size_t)I->NF * sizeof(double), hipMemcpyDeviceToHost));
  return WQL_OK;
}

extern "C" int wql_iface_time_step(wql_iface* I, double dt, int32_t nsteps, double* t) {
  if (int rc = iface_ok(I)) return rc;
  if (!t) return fail(WQL_EINVAL, "null time");
  for (int n = 0; n < nsteps; ++n) {
    const double t0 = *t;
    for (int s = 0; s < 5; ++s) {
      if (s == 0) {
        int rc = record(I->L);
        if (!rc) rc = record(I->R);
        if (rc) return rc;
      }
      // the blocks' last-stage rates are dead when another step follows
      // (as in wql_time_step)
      const bool dead = s == 4 && n + 1 < nsteps && kRkA[0] == 0.0;
      int rc = iface_stage(I, kRkA[s], kRkB[s] * dt, *t, s == 0, dead);
      if (rc) return rc;
      *t = t0 + kRkC[s + 1] * dt;
    }
  }
  return WQL_OK;
}

static int iface_sync(wql_iface* I) {
  HIP_TRY(hipStreamSynchronize(I->L->stream));
  HIP_TRY(hipStreamSynchronize(I->R->stream));
  return WQL_OK;
}

extern "C" int wql_iface_set_frames(wql_iface* I, int32_t nslots) {
  if (int rc = iface_ok(I)) return rc;
  if (nslots < 0) return fail(WQL_EINVAL, "nslots must be >= 0");
  if (int rc = iface_sync(I)) return rc;
  if (I->frames) (void)hipFree(I->frames);
  I->frames = nullptr;
  I->nslots = 0;
  I->slot = 0;
  if (nslots == 0) return WQL_OK;
  const size_t bytes = (size_t)nslots * kFramePlanes * I->NF * sizeof(double);
  hipError_t e = hipMalloc((void**)&I->frames, bytes);
  if (e != hipSuccess) return fail(WQL_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  HIP_TRY(launch_fill<double>(I->frames, (long long)nslots * kFramePlanes * I->NF, 0.0, I->L->stream));
  HIP_TRY(hipStreamSynchronize(I->L->stream));
  I->nslots = nslots;
  return WQL_OK;
}

extern "C" int wql_iface_get_frames(wql_iface* I, int32_t nslots, double* out) {
  if (int rc = iface_ok(I)) return rc;
  if (!out || nslots < 0 || nslots > I->nslots) return fail(WQL_EINVAL, "bad slot count or buffer");
  if (int rc = iface_sync(I)) return rc;
  if (nslots)
    HIP_TRY(hipMemcpy(out, I->frames, (size_t)nslots * kFramePlanes * I->NF * sizeof(double),
                      hipMemcpyDeviceToHost));
  return WQL_OK;
}

extern "C" int wql_iface_get_state(wql_iface* I, double* S, double* DS, double* trup) {
  if (int rc = iface_ok(I)) return rc;
  if (int rc = iface_sync(I)) return rc;
  const size_t nb = (size_t)I->NF * sizeof(double);
  if (S) HIP_TRY(hipMemcpy(S, I->S, 4 * nb, hipMemcpyDeviceToHost));
  if (DS) HIP_TRY(hipMemcpy(DS, I->DS, 4 * nb, hipMemcpyDeviceToHost));
  if (trup) HIP_TRY(hipMemcpy(trup, I->trup, nb, hipMemcpyDeviceToHost));
  return WQL_OK;
}

extern "C" int wql_grid_metrics(int32_t nq, int32_t nr, int32_t ns, const double* X, double* mx,
                                double* my, double* mz, double* J) {
  if (nq < 12 || nr < 12 || ns < 12) return fail(WQL_EINVAL, "SBP-6 needs >= 12 nodes per axis");
  if (!X || !mx || !my || !mz || !J) return fail(WQL_EINVAL, "null array");
  HIP_TRY(launch_metrics(nq, nr, ns, X, mx, my, mz, J, 0));
  HIP_TRY(hipStreamSynchronize(0));
  return WQL_OK;
}

// ---------------------------------------------------------------------------
// discrete adjoint of the fused stage (wql3d_adjoint.hip)
// ---------------------------------------------------------------------------
static void adjoint_free(wql_block* b) {
  void* ptrs[] = {b->adj_Fb, b->adj_Eb, b->adj_DFb, b->adj_DEb, b->adj_g, b->adj_Z, b->slot_F,
                  b->slot_E};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b->adj_Fb = b->adj_DFb = b->adj_g = b->slot_F = nullptr;
  b->adj_Eb = b->adj_DEb = b->adj_Z = b->slot_E = nullptr;
  free(b->slot_mz);
  b->slot_mz = nullptr;
  b->adj_nslots = 0;
  b->adj_nslots_f = 0;
}

// bytes of one checkpoint of F: the RK state, and under FP32-Kahan its
// compensation as well (a reloaded state then steps bit-identically)
static size_t slot_F_bytes(const wql_block* b) {
  const size_t n = 9 * (size_t)b->P;
  return n * b->ssz + (b->d.precision == WQL_FP32_KAHAN ? n * sizeof(float) : 0);
}

extern "C" int wql_adjoint_create(wql_block* b, int32_t nslots) {
  return wql_adjoint_create_f(b, nslots, 0);
}

extern "C" int wql_adjoint_create_f(wql_block* b, int32_t nslots, int32_t nslots_f) {
  if (!b) return fail(WQL_EINVAL, "null block");
  if (nslots < 0 || nslots_f < 0) return fail(WQL_EINVAL, "slot counts must be >= 0");
  for (int f = 0; f < 6; ++f)
    if (b->d.pml[f]) return fail(WQL_EINVAL, "the adjoint does not differentiate PML slabs (face %d)", f);
  if (b->iface) return fail(WQL_EINVAL, "the adjoint does not differentiate interfaces");
  if (b->d.sat_form != WQL_SAT_BC)
    return fail(WQL_EINVAL, "the adjoint differentiates the BCm/BCp SAT form only");
  HIP_TRY(hipStreamSynchronize(b->stream));
  adjoint_free(b);
  const long long P = b->P, ne = 6LL * b->d.nmech * P;
  const size_t es = b->esz;
  int rc = 0;
  if (!rc) rc = dalloc(b, (void**)&b->adj_Fb, 9 * P * sizeof(double));
  if (!rc) rc = dalloc(b, &b->adj_Eb, ne * es);
  if (!rc) rc = dalloc(b, (void**)&b->adj_DFb, 9 * P * sizeof(double));
  if (!rc) rc = dalloc(b, &b->adj_DEb, ne * es);
  if (!rc) rc = dalloc(b, (void**)&b->adj_g, 3 * P * sizeof(double));
  if (!rc) rc = dalloc(b, &b->adj_Z, 18 * P * es);
  if (!rc) rc = dalloc(b, (void**)&b->slot_F, ((size_t)nslots + (size_t)nslots_f) * slot_F_bytes(b));
  if (!rc) rc = dalloc(b, &b->slot_E, (size_t)nslots * ne * es);
  if (!rc) {
    b->slot_mz = (int*)calloc(nslots > 0 ? nslots : 1, sizeof(int));
    if (!b->slot_mz) rc = fail(WQL_ENOMEM, "host allocation failed");
  }
  if (rc) { adjoint_free(b); return rc; }
  b->adj_nslots = nslots;
  b->adj_nslots_f = nslots_f;
  return wql_adjoint_reset(b);
}

extern "C" int wql_adjoint_reset(wql_block* b) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  const long long P = b->P, ne = 6LL * b->d.nmech * P;
  HIP_TRY(launch_fill<double>(b->adj_Fb, 9 * P, 0.0, b->stream));
  HIP_TRY(launch_fill<double>(b->adj_DFb, 9 * P, 0.0, b->stream));
  HIP_TRY(launch_fill<double>(b->adj_g, 3 * P, 0.0, b->stream));
  if (ne) {
    if (b->d.precision == WQL_FP64) {
      HIP_TRY(launch_fill<double>((double*)b->adj_Eb, ne, 0.0, b->stream));
      HIP_TRY(launch_fill<double>((double*)b->adj_DEb, ne, 0.0, b->stream));
    } else {
      HIP_TRY(launch_fill<float>((float*)b->adj_Eb, ne, 0.0, b->stream));
      HIP_TRY(launch_fill<float>((float*)b->adj_DEb, ne, 0.0, b->stream));
    }
  }
  return WQL_OK;
}

// slots 0 .. adj_nslots-1 hold F and eta, the adj_nslots_f after them F only
static int check_slot(wql_block* b, int slot, uint32_t mask) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  const int n = b->adj_nslots + b->adj_nslots_f;
  if (slot < 0 || slot >= n) return fail(WQL_EINVAL, "slot %d outside 0..%d", slot, n - 1);
  if ((mask & WQL_MEMORY) && slot >= b->adj_nslots)
    return fail(WQL_EINVAL, "slot %d holds F only (memory variables: slots 0..%d)", slot,
                b->adj_nslots - 1);
  return WQL_OK;
}

extern "C" int wql_state_save_part(wql_block* b, int32_t slot, uint32_t mask) {
  if (mask & ~(uint32_t)(WQL_F | WQL_MEMORY)) return fail(WQL_EINVAL, "mask: WQL_F and/or WQL_MEMORY");
  int rc = check_slot(b, slot, mask);
  if (rc) return rc;
  const long long P = b->P, ne = 6LL * b->d.nmech * P;
  char* sf = (char*)b->slot_F + (size_t)slot * slot_F_bytes(b);
  if (mask & WQL_F) {
    HIP_TRY(hipMemcpyAsync(sf, b->F[b->cur], 9 * P * b->ssz, hipMemcpyDeviceToDevice, b->stream));
    if (b->d.precision == WQL_FP32_KAHAN)
      HIP_TRY(hipMemcpyAsync(sf + 9 * P * b->ssz, b->Fc, 9 * P * sizeof(float),
                             hipMemcpyDeviceToDevice, b->stream));
  }
  if (mask & WQL_MEMORY) {
    if (ne)
      HIP_TRY(hipMemcpyAsync((char*)b->slot_E + (size_t)slot * ne * b->esz, b->eta, ne * b->esz,
                             hipMemcpyDeviceToDevice, b->stream));
    b->slot_mz[slot] = b->mech_zero;
  }
  return WQL_OK;
}

extern "C" int wql_state_load_part(wql_block* b, int32_t slot, uint32_t mask) {
  if (mask & ~(uint32_t)(WQL_F | WQL_MEMORY)) return fail(WQL_EINVAL, "mask: WQL_F and/or WQL_MEMORY");
  int rc = check_slot(b, slot, mask);
  if (rc) return rc;
  const long long P = b->P, ne = 6LL * b->d.nmech * P;
  const char* sf = (const char*)b->slot_F + (size_t)slot * slot_F_bytes(b);
  if (mask & WQL_F) {
    b->fo_valid = false;
    HIP_TRY(hipMemcpyAsync(b->F[b->cur], sf, 9 * P * b->ssz, hipMemcpyDeviceToDevice, b->stream));
    if (b->d.precision == WQL_FP32_KAHAN)
      HIP_TRY(hipMemcpyAsync(b->Fc, sf + 9 * P * b->ssz, 9 * P * sizeof(float),
                             hipMemcpyDeviceToDevice, b->stream));
  }
  if (mask & WQL_MEMORY) {
    if (ne)
      HIP_TRY(hipMemcpyAsync(b->eta, (char*)b->slot_E + (size_t)slot * ne * b->esz, ne * b->esz,
                             hipMemcpyDeviceToDevice, b->stream));
    // the slot's zero mechanisms; Deta of those is zero as well (it only ever
    // holds A*0 + 0)
    b->mech_zero = b->slot_mz[slot];
  }
  return WQL_OK;
}

extern "C" int wql_state_save(wql_block* b, int32_t slot) {
  return wql_state_save_part(b, slot, WQL_F | WQL_MEMORY);
}

extern "C" int wql_state_load(wql_block* b, int32_t slot) {
  return wql_state_load_part(b, slot, WQL_F | WQL_MEMORY);
}

template <typename T, typename St = double>
static int adjoint_stage_t(wql_block* b, double A, double Bdt) {
  if (int rc = ensure_operand_copy(b)) return rc;   // the gradient differences F
  StepArgs<T, St> a = make_args<T, St>(b, 0.0);
  AdjArgs<T> j;
  j.Fb = b->adj_Fb;
  j.Eb = (T*)b->adj_Eb;
  j.DFb = b->adj_DFb;
  j.DEb = (T*)b->adj_DEb;
  for (int k = 0; k < 3; ++k) j.g[k] = b->adj_g + k * b->P;
  j.Z = (T*)b->adj_Z;
  j.xcd = b->tune_adj_xcd;
  HIP_TRY((launch_adjoint_stage<T, St>(a, j, b->d.mesh, A, Bdt, b->stream)));
  return WQL_OK;
}

extern "C" int wql_adjoint_stage(wql_block* b, double A, double Bdt) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  switch (b->d.precision) {
    case WQL_FP64: return adjoint_stage_t<double>(b, A, Bdt);
    case WQL_FP32_MIXED: return adjoint_stage_t<float>(b, A, Bdt);
    default: return adjoint_stage_t<float, float>(b, A, Bdt);
  }
}

extern "C" int wql_adjoint_add_receivers(wql_block* b, const double* g) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  if (!g) return fail(WQL_EINVAL, "null array");
  if (!b->nrec) return fail(WQL_ESTATE, "no receivers (wql_block_set_receivers)");
  HIP_TRY(launch_adjoint_inject(b->adj_Fb, b->P, b->rec_idx, b->nrec, g, b->stream));
  return WQL_OK;
}

// dense device arrays (x fastest, no padding) <-> the adjoint planes
template <typename T>
static int adj_copy(wql_block* b, double* dense, void* planes, int ncomp, bool to_dense) {
  const long long P = b->P;
  const long long nodes = (long long)b->nx * b->ny * b->nz;
  for (int c = 0; c < ncomp; ++c) {
    T* pl = (T*)planes + c * P;
    double* dn = dense + c * nodes;
    if (to_dense)
      HIP_TRY(launch_scatter<T>(dn, pl, b->nx, b->ny, b->nz, 1, 1, b->nx, (long long)b->nx * b->ny,
                                0, 0, b->stream));
    else
      HIP_TRY(launch_gather<T>(pl, dn, b->nx, b->ny, b->nz, 1, 1, b->nx, (long long)b->nx * b->ny,
                               0, 0, b->stream));
  }
  return WQL_OK;
}

static int adj_state_io(wql_block* b, double* Fbar, double* etabar, bool to_dense) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  if (!Fbar) return fail(WQL_EINVAL, "null array");
  int rc = adj_copy<double>(b, Fbar, b->adj_Fb, 9, to_dense);
  if (rc) return rc;
  if (b->d.nmech) {
    if (!etabar) return fail(WQL_EINVAL, "null eta array");
    rc = b->d.precision == WQL_FP64
             ? adj_copy<double>(b, etabar, b->adj_Eb, 6 * b->d.nmech, to_dense)
             : adj_copy<float>(b, etabar, b->adj_Eb, 6 * b->d.nmech, to_dense);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return WQL_OK;
}

extern "C" int wql_adjoint_get_state(wql_block* b, double* Fbar, double* etabar) {
  return adj_state_io(b, Fbar, etabar, true);
}

extern "C" int wql_adjoint_set_state(wql_block* b, const double* Fbar, const double* etabar) {
  return adj_state_io(b, const_cast<double*>(Fbar), const_cast<double*>(etabar), false);
}

extern "C" int wql_adjoint_gradient(wql_block* b, double* g_rho, double* g_lambda, double* g_mu) {
  if (!b || !b->adj_Fb) return fail(WQL_ESTATE, "no adjoint context (wql_adjoint_create)");
  if (!g_rho || !g_lambda || !g_mu) return fail(WQL_EINVAL, "null array");
  double* out[3] = {g_rho, g_lambda, g_mu};
  for (int k = 0; k < 3; ++k) {
    int rc = adj_copy<double>(b, out[k], b->adj_g + k * b->P, 1, true);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return WQL_OK;
}
