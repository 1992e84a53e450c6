// fv_capi.hip -- the C-ABI of include/amrvac_fv.h: device-resident level
// state, kernel selection, and the auxiliary device kernels (CFL dt
// reduction, volume averages, ghost-cell fill for download).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "amrvac_fv.h"
#include "fv_launch.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(AMRVAC_FV_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

constexpr double kBig = 1.0e99;     // src/mod_constants.t:11
constexpr double kSmall = 1.0e-12;  // src/mod_constants.t:8

// Scratch of the allocation-free dt reduction (amrvac_fv_getdt_state_noalloc):
// per-block partial maxima, 2 doubles per block for at most kNoallocMaxBlocks
// blocks, then the reduced triple.  Every GPU the library runs on holds its
// own copy (512 KiB) until the process ends.  In host code the array's name
// is not a device address: kernels get the address hipGetSymbolAddress
// returned for the context's device (amrvac_fv_ctx::scratch).
constexpr int kNoallocMaxBlocks = 32768;
constexpr int kMaxDevices = 64;
}  // namespace
static __device__ __align__(256) double scratch[2 * kNoallocMaxBlocks + 4];
namespace {
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the device address of `scratch` on the current device, looked up once
int scratch_address(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(AMRVAC_FV_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(AMRVAC_FV_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return AMRVAC_FV_OK;
}

}  // namespace

struct amrvac_fv_ctx {
  amrvac_fv_config cfg{};
  int nw = 0;
  int nnb = 0;
  int nb = 0;
  int bx = 16;       // block_nx(1) = block_nx(2)
  int bnz = 1;
  size_t cells = 0;  // interior cells per block
  size_t cells_g = 0;  // ghosted cells per block
  int nxg[3] = {1, 1, 1};
  double dx[3] = {1, 1, 1};
  int bc[6] = {5, 5, 5, 5, 5, 5};
  double* slot[4] = {nullptr, nullptr, nullptr, nullptr};  // ps, ps1, ps2, ps3
  // constrained transport: face field of each slot [nb][ndim][cells] (the
  // upper faces of each cell), the per-stage face buffer, the new faces
  int nws = 0;
  double* wsl[4] = {nullptr, nullptr, nullptr, nullptr};
  double* d_ctf = nullptr;
  double* d_wsn = nullptr;
  int* d_nbr = nullptr;
  int* d_crash = nullptr;
  double* d_red = nullptr;    // per-block reduction results
  // CFL partials of ps left by the last stage that wrote it (uniform level,
  // production precision): getdt of ps then only reduces them
  double* d_dtp = nullptr;
  bool dt_valid = false;
  bool slot0_exposed = false;  // ps handed out by amrvac_fv_slot_pointer
  double* d_red2 = nullptr;   // final reduction
  double* d_stage = nullptr;  // ghosted staging for download
  std::vector<double> h_stage;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  fvc::FVParams fp{};
  // block-structured AMR: virtual blocks after the nb leaves hold the ghost
  // cells across level jumps (restricted from finer or prolonged from
  // coarser neighbours); the sweep's neighbour table points at them
  bool amr = false;
  bool fix_conserve = false;
  int nv = 0;                // virtual blocks
  int nface = 0;             // register stride (transverse cells of a face)
  int* d_level = nullptr;    // [nb] level-1
  int* d_ntype = nullptr;    // [nb][nnb]
  int* d_nchild = nullptr;   // [nb][nnb][nch]
  int4* d_rtask = nullptr;   // restriction: (coarse block, direction, virtual block, -)
  int4* d_ptask = nullptr;   // prolongation: (fine block, direction, virtual block, coarse block)
  int2* d_ptask2 = nullptr;  // prolongation: (packed position in parent, packed direction coarse->parent)
  int* d_ftask = nullptr;    // refluxing: coarse blocks with a finer face neighbour
  int* d_fregoff = nullptr;  // [nb][2*ndim]
  double* d_freg = nullptr;  // flux registers
  double* d_xlo = nullptr;   // [nb][ndim] block lower corners
  double* d_adjG = nullptr;  // 3-D adjoint: summed z-face contributions per cell [nw][nb*cells]
  // device regrid: new-leaf state, plan, Löhner flags and weights (kept across
  // regrids, grown on demand)
  // element capacities of the level's buffers: an AMR regrid keeps every
  // buffer that still fits the new level (no free / allocate per regrid)
  size_t slot_cap = 0, nbr_cap = 0, red_cap = 0, dtp_cap = 0, stage_cap = 0;
  int nreg = 0;  // refluxing registers of the level
  size_t level_cap = 0, ntype_cap = 0, nchild_cap = 0, rtask_cap = 0, ptask_cap = 0, ptask2_cap = 0;
  size_t ftask_cap = 0, fregoff_cap = 0, freg_cap = 0, xlo_cap = 0;
  double* d_remap = nullptr;
  size_t remap_cap = 0;
  int* d_plan = nullptr;
  size_t plan_cap = 0;
  unsigned char* d_flags = nullptr;
  size_t flags_cap = 0;
  int* d_iws = nullptr;
  double* d_wts = nullptr;
  std::vector<int> h_iws;     // the estimator's variables and weights as
  std::vector<double> h_wts;  // last uploaded
  // pinned host buffer of the results read back: the dt reduction's three
  // doubles at 0, the small-value flag at kPinCrash, the regrid flags from
  // kPinFlags followed (4-byte aligned) by the small-value flag copied with
  // them (copies into it need no staging and share one synchronisation)
  char* h_pin = nullptr;
  size_t pin_cap = 0;
  int nrt = 0, npt = 0, nft = 0;
  std::vector<int> h_level;
  double* scratch = nullptr; // this device's copy of `scratch` (no-allocation dt)
  int crash_seen = 0;        // small-value flag copied with the last dt
  bool crash_fresh = false;  // no stage launched since that copy
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  size_t nev = 0;
};

namespace fv {

template <class P>
__global__ void __launch_bounds__(NTHREADS) getdt_kernel(const double* __restrict__ w, int nblocks, int bx, int bnz,
                                                         FVParams fp, const int* __restrict__ blevel,
                                                         double* out, int* crash) {
  const int BXr = bx;  // block_nx(1) = block_nx(2)
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  if (blevel) {  // the block's level: 1/dx scales by 2^level exactly
    const double sc = double(1 << blevel[b]);
    for (int d = 0; d < 3; ++d) fp.dxinv[d] = fp.dxinv[d] * sc;
  }
  const size_t cells = (size_t)BXr * BXr * bnz;
  const double* wb = w + (size_t)b * P::NW * cells;
  double maxtot = -1.0e99, cmaxb = 0.0;
  int small = 0;
  for (size_t c = threadIdx.x; c < cells; c += NTHREADS) {
    double cons[P::NW];
#pragma unroll
    for (int iw = 0; iw < P::NW; ++iw) cons[iw] = wb[iw * cells + c];
    // getdt_courant converts with phys_to_primitive (src/mod_dt.t:195),
    // which checks the cell with fix_small_values
    if (fp.fix_small && small_cons<P>(cons, fp)) small = 1;
    // getdt_courant, maxsum, slab_uniform (src/mod_dt.t:203-221)
    cfl_cell<P>(cons, fp, maxtot, cmaxb);
  }
  if (small) *crash = 1;
  // wave64 shuffle max, then across the 4 waves through LDS
  for (int off = 32; off > 0; off >>= 1) {
    maxtot = fmax(maxtot, __shfl_xor(maxtot, off, 64));
    cmaxb = fmax(cmaxb, __shfl_xor(cmaxb, off, 64));
  }
  __shared__ double s[2][NTHREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { s[0][wv] = maxtot; s[1][wv] = cmaxb; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s[0][0], c = s[1][0];
    for (int i = 1; i < NTHREADS / 64; ++i) { m = fmax(m, s[0][i]); c = fmax(c, s[1][i]); }
    out[2 * b] = m;
    out[2 * b + 1] = c;
  }
}

// out[0] = max of the per-block maxtot, out[1] = max cmax, out[2] = the
// context's small-value flag (it reaches the host with the dt copy)
__global__ void __launch_bounds__(NTHREADS) reduce_max2_kernel(const double* __restrict__ in, int n,
                                                               double* out, const int* __restrict__ crash) {
  double m = -1.0e99, c = 0.0;
  for (int i = threadIdx.x; i < n; i += NTHREADS) {
    m = fmax(m, in[2 * i]);
    c = fmax(c, in[2 * i + 1]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_xor(m, off, 64));
    c = fmax(c, __shfl_xor(c, off, 64));
  }
  __shared__ double s[2][NTHREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { s[0][wv] = m; s[1][wv] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < NTHREADS / 64; ++i) { s[0][0] = fmax(s[0][0], s[0][i]); s[1][0] = fmax(s[1][0], s[1][i]); }
    out[0] = s[0][0];
    out[1] = s[1][0];
    out[2] = *crash ? 1.0 : 0.0;
  }
}

// per-block sums of w and w^2 over the interior, [nb][2][nw]
// State combination between stages (the update loops of advect,
// src/mod_advance.t:113-133, 269-291, 485-507, 515-557): dst = a postfix
// program over the slots, one rounding per operation, elementwise over the
// leaves' interior cells.  The program is uniform over the grid and checked
// on the host (depth <= COMB_DEPTH, constants in range).
constexpr int COMB_OPS = 32, COMB_CONSTS = 16, COMB_DEPTH = 8;
enum { COP_CONST = 8, COP_ADD = 9, COP_SUB = 10, COP_MUL = 11, COP_DIV = 12, COP_NEG = 13 };
struct CombProg {
  int nops;
  int ops[COMB_OPS];
  double k[COMB_CONSTS];
};

__global__ void __launch_bounds__(NTHREADS) combine_kernel(double* dst, const double* s0, const double* s1,
                                                           const double* s2, const double* s3, size_t n,
                                                           CombProg p) {
  const size_t stride = (size_t)gridDim.x * NTHREADS;
  for (size_t e = (size_t)blockIdx.x * NTHREADS + threadIdx.x; e < n; e += stride) {
    double st[COMB_DEPTH];
    int sp = 0, kc = 0;
    for (int q = 0; q < p.nops; ++q) {
      const int op = p.ops[q];
      if (op <= 3) {
        st[sp++] = op == 0 ? s0[e] : op == 1 ? s1[e] : op == 2 ? s2[e] : s3[e];
      } else if (op == COP_CONST) {
        st[sp++] = p.k[kc++];
      } else if (op == COP_NEG) {
        st[sp - 1] = -st[sp - 1];
      } else {
        const double a = st[sp - 2], b = st[sp - 1];
        --sp;
        st[sp - 1] = op == COP_ADD ? a + b : op == COP_SUB ? a - b : op == COP_MUL ? a * b : a / b;
      }
    }
    dst[e] = st[0];
  }
}

__global__ void __launch_bounds__(NTHREADS) volsum_kernel(const double* __restrict__ w, int nblocks, int nw,
                                                          size_t cells, double* out) {
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  __shared__ double s[NTHREADS / 64];
  for (int iw = 0; iw < nw; ++iw) {
    const double* q = w + ((size_t)b * nw + iw) * cells;
    for (int p = 1; p <= 2; ++p) {
      double acc = 0.0;
      for (size_t c = threadIdx.x; c < cells; c += NTHREADS) {
        const double v = q[c];
        acc += p == 1 ? v : v * v;
      }
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
      if (lane == 0) s[wv] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < NTHREADS / 64; ++i) t += s[i];
        out[((size_t)b * 2 + (p - 1)) * nw + iw] = t;
      }
      __syncthreads();
    }
  }
}

// ghost-cell fill for download: the ghosted reference layout of every block,
// same-level copies and physical boundaries as getbc
// (src/mod_ghostcells_update.t:350-566, src/mod_boundary_conditions.t:12-120)
template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) fill_ghosted_kernel(SweepArgs a, int nw, int ng, int nxg0, int nxg1,
                                                                int nxg2, double* out) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  const int b = blockIdx.y;
  const size_t cg = (size_t)nxg0 * nxg1 * nxg2;
  const size_t cells = (size_t)BXr * BXr * a.bnz;
  for (size_t c = blockIdx.x * (size_t)NTHREADS + threadIdx.x; c < cg; c += (size_t)gridDim.x * NTHREADS) {
    int i = int(c % nxg0);
    int j = int((c / nxg0) % nxg1);
    int k = int(c / ((size_t)nxg0 * nxg1));
    int x = i - ng, y = NDIM > 1 ? j - ng : 0, z = NDIM > 2 ? k - ng : 0;
    double sign;
    const int sb = locate<NDIM>(a, b, x, y, z, sign);
    const size_t so = (size_t)sb * nw * cells + ((size_t)z * BXr + y) * BXr + x;
    for (int iw = 0; iw < nw; ++iw) out[((size_t)b * nw + iw) * cg + c] = sign * a.src[so + iw * cells];
  }
}


// ---------------------------------------------------------------------------
// block-structured AMR: ghost cells across level jumps and refluxing
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode_dir(int k, int ndim, int* i) {
  for (int d = 0; d < 3; ++d) {
    if (d < ndim) { i[d] = k % 3 - 1; k /= 3; } else i[d] = 0;
  }
}

// Ghost cells of a coarse block facing a finer neighbour: the mean of the
// 2^ndim fine cells each covers, summed in array-element order
// (coarsen_grid src/amr/mod_coarsen.t:12-79 into the coarse representation,
// bc_fill_restrict src/mod_ghostcells_update.t:855-898).  Written into the
// direction's virtual block in the neighbour's frame.
template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) restrict_kernel(double* w, int nw, int bx, int bnz, int ng, int nnb,
                                                            const int4* task, const int* nchild) {
  const int BXr = bx;  // block_nx(1) = block_nx(2)
  const int4 t = task[blockIdx.x];
  const int C = t.x, k = t.y, V = t.z;
  int i[3];
  decode_dir(k, NDIM, i);
  const int n[3] = {BXr, BXr, bnz};
  int lo[3] = {0, 0, 0}, cnt[3] = {1, 1, 1};
  for (int d = 0; d < NDIM; ++d) {
    lo[d] = i[d] == -1 ? -ng : (i[d] == 0 ? 0 : n[d]);
    cnt[d] = i[d] == 0 ? n[d] : ng;
  }
  const size_t cells = (size_t)BXr * BXr * bnz;
  const double cofi = 1.0 / double(1 << NDIM);
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int c = threadIdx.x; c < total; c += NTHREADS) {
    const int g[3] = {lo[0] + c % cnt[0], lo[1] + (c / cnt[0]) % cnt[1], lo[2] + c / (cnt[0] * cnt[1])};
    int gn[3] = {0, 0, 0}, f[3] = {0, 0, 0}, cidx = 0;
    for (int d = 0; d < NDIM; ++d) {
      gn[d] = g[d] - i[d] * n[d];
      const int ic = gn[d] >= n[d] / 2 ? 2 : 1;
      cidx += (ic - 1) << d;
      f[d] = 2 * (gn[d] - (ic - 1) * (n[d] / 2));
    }
    const int F = nchild[((size_t)C * nnb + k) * (1 << NDIM) + cidx];
    const double* wf = w + (size_t)F * nw * cells;
    double* wv = w + (size_t)V * nw * cells + ((size_t)gn[2] * BXr + gn[1]) * BXr + gn[0];
    for (int iw = 0; iw < nw; ++iw) {
      double sum = 0.0;
      bool first = true;
      for (int dk = 0; dk <= (NDIM > 2 ? 1 : 0); ++dk)
        for (int dj = 0; dj <= 1; ++dj)
          for (int di = 0; di <= 1; ++di) {
            const double v = wf[iw * cells + ((size_t)(f[2] + dk) * BXr + (f[1] + dj)) * BXr + f[0] + di];
            sum = first ? v : sum + v;
            first = false;
          }
      wv[iw * cells] = sum * cofi;
    }
  }
}

// one cell of the coarse representation of fine block F (1-based index p of
// the (n/2 + 2ng)^ndim coarse grid around F), variable iw, as bc_prolong
// sees it: physical boundary rule on the coarse grid where F touches the
// boundary (bc_phys), the mean of F's own fine cells inside F, and the
// coarse neighbour's data (its interior or, through its neighbour table,
// its same-level or restricted ghost cells) elsewhere
template <int NDIM>
__device__ double psc_value(const SweepArgs& a, int nw, int ng, int F, int C, const int* icF, const int* iC,
                            const bool (*bnd)[2], int p0, int p1, int p2, int iw) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  const int n[3] = {BXr, BXr, a.bnz};
  int q[3] = {p0, p1, p2};
  double sign = 1.0;
  bool inside = true;
  for (int d = 0; d < NDIM; ++d) {
    const int clo = ng + 1, chi = ng + n[d] / 2;
    if (q[d] < clo && bnd[d][0]) {
      const int tp = a.bc[2 * d];
      if (tp == 2) q[d] = clo;
      else { q[d] = 2 * clo - 1 - q[d]; if (tp == 4) sign = -sign; }
    } else if (q[d] > chi && bnd[d][1]) {
      const int tp = a.bc[2 * d + 1];
      if (tp == 2) q[d] = chi;
      else { q[d] = 2 * chi + 1 - q[d]; if (tp == 4) sign = -sign; }
    }
    if (q[d] < clo || q[d] > chi) inside = false;
  }
  const size_t cells = (size_t)BXr * BXr * a.bnz;
  if (inside) {
    int f[3] = {0, 0, 0};
    for (int d = 0; d < NDIM; ++d) f[d] = 2 * (q[d] - (ng + 1));
    const double* wf = a.src + (size_t)F * nw * cells + iw * cells;
    double sum = 0.0;
    bool first = true;
    for (int dk = 0; dk <= (NDIM > 2 ? 1 : 0); ++dk)
      for (int dj = 0; dj <= 1; ++dj)
        for (int di = 0; di <= 1; ++di) {
          const double v = wf[((size_t)(f[2] + dk) * BXr + (f[1] + dj)) * BXr + f[0] + di];
          sum = first ? v : sum + v;
          first = false;
        }
    return sign * (sum * (1.0 / double(1 << NDIM)));
  }
  int x[3] = {0, 0, 0};
  for (int d = 0; d < NDIM; ++d) x[d] = q[d] + (icF[d] - 1) * (n[d] / 2) + n[d] * iC[d] - ng - 1;
  double s2;
  const int sb = locate<NDIM>(a, C, x[0], x[1], x[2], s2);
  return sign * s2 * a.src[(size_t)sb * nw * cells + iw * cells + ((size_t)x[2] * BXr + x[1]) * BXr + x[0]];
}

// Ghost cells of a fine block facing a coarser neighbour: interpolation_linear
// (src/mod_ghostcells_update.t:1381-1525, slab_uniform, conserved variables)
// from the coarse representation with limited slopes, as bc_prolong
// (:1261-1339) and bc_fill_prolong (:1103-1154) assemble it.
template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) prolong_kernel(SweepArgs a, double* w, int nw, int ng, int nnb,
                                                           const int4* task, const int2* task2, const int* ntype,
                                                           const int* level, const double* xlo, FVParams fp) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  const int4 t = task[blockIdx.x];
  const int2 t2 = task2[blockIdx.x];
  const int F = t.x, k = t.y, V = t.z, C = t.w;
  int i[3], icF[3] = {1, 1, 1}, iC[3] = {0, 0, 0};
  decode_dir(k, NDIM, i);
  for (int d = 0; d < NDIM; ++d) {
    icF[d] = (t2.x >> (2 * d)) & 3;
    iC[d] = ((t2.y >> (2 * d)) & 3) - 1;
  }
  const int n[3] = {BXr, BXr, a.bnz};
  bool bnd[3][2] = {{false, false}, {false, false}, {false, false}};
  for (int d = 0; d < NDIM; ++d)
    for (int sd = 0; sd < 2; ++sd) {
      int v[3] = {0, 0, 0};
      v[d] = sd == 0 ? -1 : 1;
      int kk = 0, m = 1;
      for (int e = 0; e < NDIM; ++e) { kk += (v[e] + 1) * m; m *= 3; }
      bnd[d][sd] = ntype[(size_t)F * nnb + kk] == 1;
    }
  int lo[3] = {0, 0, 0}, cnt[3] = {1, 1, 1};
  double dxFi[3] = {1, 1, 1}, dxCo[3] = {1, 1, 1}, invdxCo[3] = {1, 1, 1}, xFimin[3] = {0, 0, 0},
         xComin[3] = {0, 0, 0};
  const double sc = double(1 << level[F]);
  for (int d = 0; d < NDIM; ++d) {
    lo[d] = i[d] == -1 ? -ng : (i[d] == 0 ? 0 : n[d]);
    cnt[d] = i[d] == 0 ? n[d] : ng;
    dxFi[d] = fp.dx[d] / sc;
    dxCo[d] = 2.0 * dxFi[d];
    invdxCo[d] = 1.0 / dxCo[d];
    const double rpxmin = xlo[(size_t)F * NDIM + d];
    xFimin[d] = rpxmin - double(ng) * dxFi[d];
    xComin[d] = rpxmin - double(ng) * dxCo[d];
  }
  const size_t cells = (size_t)BXr * BXr * a.bnz;
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int c = threadIdx.x; c < total; c += NTHREADS) {
    const int g[3] = {lo[0] + c % cnt[0], lo[1] + (c / cnt[0]) % cnt[1], lo[2] + c / (cnt[0] * cnt[1])};
    int co[3] = {1, 1, 1};
    double eta[3] = {0, 0, 0};
    for (int d = 0; d < NDIM; ++d) {
      const double xFi = xFimin[d] + (double(g[d] + ng + 1) - 0.5) * dxFi[d];
      co[d] = int((xFi - xComin[d]) * invdxCo[d]) + 1;
      const double xCo = xComin[d] + (double(co[d]) - 0.5) * dxCo[d];
      eta[d] = (xFi - xCo) * invdxCo[d];
    }
    int gn[3] = {0, 0, 0};
    for (int d = 0; d < NDIM; ++d) gn[d] = g[d] - i[d] * n[d];
    double* wv = w + (size_t)V * nw * cells + ((size_t)gn[2] * BXr + gn[1]) * BXr + gn[0];
    for (int iw = 0; iw < nw; ++iw) {
      const double c0 = psc_value<NDIM>(a, nw, ng, F, C, icF, iC, bnd, co[0], co[1], co[2], iw);
      double val = c0;
      for (int d = 0; d < NDIM; ++d) {
        int h[3] = {co[0], co[1], co[2]}, p[3] = {co[0], co[1], co[2]};
        h[d] -= 1;
        p[d] += 1;
        const double slopeL = c0 - psc_value<NDIM>(a, nw, ng, F, C, icF, iC, bnd, h[0], h[1], h[2], iw);
        const double slopeR = psc_value<NDIM>(a, nw, ng, F, C, icF, iC, bnd, p[0], p[1], p[2], iw) - c0;
        const double slopeC = 0.5 * (slopeR + slopeL);
        const double signC = copysign(1.0, slopeC);
        const double m1 = fabs(slopeC) < signC * slopeL ? fabs(slopeC) : signC * slopeL;
        const double m2 = m1 < signC * slopeR ? m1 : signC * slopeR;
        const double slope = signC * (0.0 > m2 ? 0.0 : m2);
        val = val + slope * eta[d];
      }
      wv[iw * cells] = val;
    }
  }
}

// fix_conserve (src/mod_fix_conserve.t:465-595) with the registers of
// store_flux (:597-648): a coarse block's cells next to a finer neighbour
// drop their own face flux and take the sum of the fine fluxes
template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) fixcons_kernel(double* w, int nw, int bx, int bnz, int nnb, const int* ftask,
                                                           const int* ntype, const int* nchild,
                                                           const int* fregoff, const double* freg, int nface) {
  const int BXr = bx;  // block_nx(1) = block_nx(2)
  const int C = ftask[blockIdx.x];
  const int n[3] = {BXr, BXr, bnz};
  const size_t cells = (size_t)BXr * BXr * bnz;
  const double cofi = 1.0 / double(1 << NDIM);
  double* wc = w + (size_t)C * nw * cells;
  for (int idims = 0; idims < NDIM; ++idims) {
    int ta = -1, tb = -1;
    for (int d = 0; d < NDIM; ++d) {
      if (d == idims) continue;
      if (ta < 0) ta = d; else tb = d;
    }
    const int na = n[ta], nbt = tb >= 0 ? n[tb] : 1;
    for (int iside = 1; iside <= 2; ++iside) {
      int v[3] = {0, 0, 0};
      v[idims] = 2 * iside - 3;
      int k = 0, m = 1;
      for (int e = 0; e < NDIM; ++e) { k += (v[e] + 1) * m; m *= 3; }
      if (ntype[(size_t)C * nnb + k] != 4) continue;
      const int ix = iside == 1 ? 0 : n[idims] - 1;
      const double* own = freg + fregoff[C * 2 * NDIM + 2 * idims + iside - 1];
      for (int t = threadIdx.x; t < nw * na * nbt; t += NTHREADS) {
        const int iw = t / (na * nbt), r = t % (na * nbt), a2 = r % na, b2 = r / na;
        int c[3] = {0, 0, 0};
        c[idims] = ix;
        c[ta] = a2;
        if (tb >= 0) c[tb] = b2;
        const double f = own[(size_t)iw * nface + a2 + (size_t)na * b2];
        const double pflux = iside == 1 ? -f : f;
        double& x = wc[iw * cells + ((size_t)c[2] * BXr + c[1]) * BXr + c[0]];
        x = x - pflux;
      }
      __syncthreads();
      const int iother = 3 - iside;
      const int nca = na / 2, ncb = tb >= 0 ? nbt / 2 : 1;
      const int nkid = 1 << (NDIM - 1);
      for (int t = threadIdx.x; t < nkid * nw * nca * ncb; t += NTHREADS) {
        const int kid = t / (nw * nca * ncb), r0 = t % (nw * nca * ncb);
        const int iw = r0 / (nca * ncb), r = r0 % (nca * ncb), a2 = r % nca, b2 = r / nca;
        int ic[3] = {1, 1, 1};
        ic[idims] = v[idims] == -1 ? 2 : 1;
        ic[ta] = 1 + (kid & 1);
        if (tb >= 0) ic[tb] = 1 + ((kid >> 1) & 1);
        int cidx = 0;
        for (int d = 0; d < NDIM; ++d) cidx += (ic[d] - 1) << d;
        const int F = nchild[((size_t)C * nnb + k) * (1 << NDIM) + cidx];
        const double* q = freg + fregoff[F * 2 * NDIM + 2 * idims + iother - 1] + (size_t)iw * nface;
        const int fa = 2 * a2, fb = tb >= 0 ? 2 * b2 : 0;
        double sum = q[fa + (size_t)na * fb];
        sum = sum + q[fa + 1 + (size_t)na * fb];
        if (tb >= 0) {
          sum = sum + q[fa + (size_t)na * (fb + 1)];
          sum = sum + q[fa + 1 + (size_t)na * (fb + 1)];
        }
        const double pf = iother == 1 ? sum : -sum;
        int c[3] = {0, 0, 0};
        c[idims] = ix;
        c[ta] = (ic[ta] - 1) * nca + a2;
        if (tb >= 0) c[tb] = (ic[tb] - 1) * ncb + b2;
        double& x = wc[iw * cells + ((size_t)c[2] * BXr + c[1]) * BXr + c[0]];
        x = x + pf * cofi;
      }
      __syncthreads();
    }
  }
}

template <class P>
hipError_t launch_getdt(const double* w, int nb, int bx, int bnz, const FVParams& fp, const int* blevel, double* out,
                        int* crash, hipStream_t s) {
  hipLaunchKernelGGL((getdt_kernel<P>), dim3(nb), dim3(NTHREADS), 0, s, w, nb, bx, bnz, fp, blevel, out, crash);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// regrid on the device: Löhner flags and the data of the new leaves
// ---------------------------------------------------------------------------
// variable iw of cell (x, y, z) of block b, ghost cells through the neighbour
// table (the values download would write into the ghosted layout)
template <int NDIM>
__device__ __forceinline__ double cell_value(const SweepArgs& a, int nw, int b, int x, int y, int z, int iw) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  double sign;
  const int sb = locate<NDIM>(a, b, x, y, z, sign);
  const size_t cells = (size_t)BXr * BXr * a.bnz;
  return sign * a.src[(size_t)sb * nw * cells + iw * cells + ((size_t)z * BXr + y) * BXr + x];
}

// add_source_linde (src/mhd/mod_mhd_phys.t:5599-5664) after the flux
// update of a stage, for every block: Linde's div B diffusion of the stage
// output dst from the stage input wCT = a.src (ghost cells through the
// neighbour table, as the stage kernel reads them):
//   g = divbdiff / (sum_d 1/dx_d^2) * (divB(x+e) - divB(x-e)) / (2 dx),
//   B_idim += g,  e += B_idim(wCT) g,
// dimension by dimension, with div B the central divvector of wCT
// (get_divb src/physics/mod_functions_bfield.t:14-45, src/mod_geometry.t:
// 582-655), and no update in the cell layer next to a face whose
// neighbour is coarser or finer (the level-jump rule, :5616-5641).  One
// workgroup per block; div B of the block and its one-cell rim in LDS.
// Built without contraction, so every build rounds like the oracle.
constexpr int kLindeLds = 7800;  // doubles: (bx+2)^2 (bnz+2) for bx <= 16, bnz <= 22
template <class P>
__global__ void __launch_bounds__(NTHREADS) linde_kernel(SweepArgs a, FVParams fp, const int* __restrict__ ntype,
                                                         int nnb, double* __restrict__ dst) {
  constexpr int NDIM = P::NDIM, NW = P::NW;
  const int b = blockIdx.x;
  if (b >= a.nblocks) return;
  if (a.blevel && a.blevel[b]) {  // cell sizes of the block's level (exact powers of two)
    const double sc = double(1 << a.blevel[b]), sc2 = sc * sc;
    for (int d = 0; d < 3; ++d) {
      fp.dx[d] = fp.dx[d] / sc;
      fp.dxinv[d] = fp.dxinv[d] * sc;
    }
    fp.invds2 = fp.invds2 * sc2;
  }
  const int bx = a.bx, bz = NDIM == 3 ? a.bnz : 1;
  const int ex = bx + 2, ez = NDIM == 3 ? bz + 2 : 1;
  __shared__ double divb[kLindeLds];
  for (int c = threadIdx.x; c < ex * ex * ez; c += NTHREADS) {
    const int x = c % ex - 1, y = (c / ex) % ex - 1, z = NDIM == 3 ? c / (ex * ex) - 1 : 0;
    double d = 0.0;
#pragma unroll
    for (int dd = 0; dd < NDIM; ++dd) {
      const int hx = dd == 0, hy = dd == 1, hz = dd == 2;
      const double bp = cell_value<NDIM>(a, NW, b, x + hx, y + hy, z + hz, P::MAG + dd);
      const double bm = cell_value<NDIM>(a, NW, b, x - hx, y - hy, z - hz, P::MAG + dd);
      d = d + 0.5 * (bp - bm) * fp.dxinv[dd];
    }
    divb[c] = d;
  }
  __syncthreads();
  int lo[3] = {0, 0, 0}, hi[3] = {bx - 1, bx - 1, bz - 1};
  if (ntype) {
    for (int d = 0; d < NDIM; ++d)
      for (int side = 0; side < 2; ++side) {
        int k = 0, m = 1;
        for (int e = 0; e < NDIM; ++e) { k += ((e == d ? (side ? 1 : -1) : 0) + 1) * m; m *= 3; }
        const int t = ntype[(size_t)b * nnb + k];
        if (t == 2 || t == 4) {
          if (side == 0) lo[d] = lo[d] + 1;
          else hi[d] = hi[d] - 1;
        }
      }
  }
  const size_t cells = (size_t)bx * bx * a.bnz;
  double* wb = dst + (size_t)b * NW * cells;
  const double* wct = a.src + (size_t)b * NW * cells;
  int crash = 0;
  for (int c = threadIdx.x; c < bx * bx * bz; c += NTHREADS) {
    const int x = c % bx, y = (c / bx) % bx, z = NDIM == 3 ? c / (bx * bx) : 0;
    if (x < lo[0] || x > hi[0] || y < lo[1] || y > hi[1] || z < lo[2] || z > hi[2]) continue;
    // the tile's layout as filled above: no z rim in 2-D
    const int l = (x + 1) + ex * ((y + 1) + ex * (NDIM == 3 ? z + 1 : 0));
#pragma unroll
    for (int idim = 0; idim < NDIM; ++idim) {
      const int st = idim == 0 ? 1 : (idim == 1 ? ex : ex * ex);
      double g = (divb[l + st] - divb[l - st]) * 0.5;  // gradient (src/mod_geometry.t:329-403)
      g = g / fp.dx[idim];
      g = g * fp.divbdiff / fp.invds2;
      wb[(P::MAG + idim) * cells + c] = wb[(P::MAG + idim) * cells + c] + g;
      wb[P::E * cells + c] = wb[P::E * cells + c] + wct[(P::MAG + idim) * cells + c] * g;
    }
    if (fp.fix_small) {  // add_source_linde checks the updated cells (:5662)
      double w[NW];
#pragma unroll
      for (int iw = 0; iw < NW; ++iw) w[iw] = wb[iw * cells + c];
      if (small_cons<P>(w, fp)) crash = 1;
    }
  }
  if (crash) *a.crash = 1;
}

template <class P>
hipError_t launch_linde(const SweepArgs& a, const FVParams& fp, const int* ntype, int nnb, double* dst,
                        hipStream_t s) {
  if constexpr (!P::MHD || P::NDIM < 2) {
    return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((linde_kernel<P>), dim3(a.nblocks), dim3(NTHREADS), 0, s, a, fp, ntype, nnb, dst);
    return hipGetLastError();
  }
}

// lohner_grid (src/amr/mod_errest.t:63-199, Cartesian, no B0, no log
// variables): per cell
//   error = sum_w weight_w sqrt(num_w / max(den_w, 1e-6)),
//   num = sum_{d,d2} (dq_d(x+e_d2) - dq_d(x-e_d2))^2,  dq_d(p) = q(p+e_d) - q(p-e_d),
//   den = sum_{d,d2} (|q(x+2e_d)-q(x)| + |q(x)-q(x-2e_d)|
//                     + wavefilter (aq_d(x+e_d2) + aq_d(x-e_d2)))^2,  aq_d(p) = |q(p+e_d)| + |q(p-e_d)|,
// evaluated in that order without contraction.  flags[b]: bit 0 when some
// cell has error >= thr, bit 1 when every cell has error < thr and
// error <= cthr (the coarsening test of errest, :10-61).
// In 2-D (16^2 blocks, one cell per lane) the weighted variables' block,
// with the cells the stencil reaches (two along each axis, one diagonally),
// is staged in LDS, up to kLohnerV variables at a time, from the same cell
// values the 3-D path reads one by one.
constexpr int kLohnerT = 16 + 4;  // 2-D tile edge: block_nx + 2 + 2
constexpr int kLohnerV = 4;
__global__ void __launch_bounds__(NTHREADS) lohner2_kernel(SweepArgs a, int nw, int nwt, const int* __restrict__ iws,
                                                           const double* __restrict__ wts, double wavefilter,
                                                           double thr, double cthr, unsigned char* __restrict__ flags,
                                                           const int* __restrict__ crash, int* __restrict__ crash_copy) {
  constexpr int T = kLohnerT, N = T - 4;
  __shared__ double tile[kLohnerV][T * T];
  const int b = blockIdx.x;
  const int x = threadIdx.x % N, y = threadIdx.x / N;
  double error = 0.0;
  for (int t0 = 0; t0 < nwt; t0 += kLohnerV) {
    const int nv = nwt - t0 < kLohnerV ? nwt - t0 : kLohnerV;
    for (int c = threadIdx.x; c < nv * T * T; c += NTHREADS) {
      const int v = c / (T * T), e = c % (T * T);
      const int X = e % T - 2, Y = e / T - 2;
      const int ox = X < 0 ? -X : (X >= N ? X - N + 1 : 0), oy = Y < 0 ? -Y : (Y >= N ? Y - N + 1 : 0);
      if ((ox == 0 || oy == 0) || (ox == 1 && oy == 1)) tile[v][e] = cell_value<2>(a, nw, b, X, Y, 0, iws[t0 + v]);
    }
    __syncthreads();
    for (int v = 0; v < nv; ++v) {
      // q at p + s1 e_d1 + s2 e_d2, as the 3-D path's q()
      auto q = [&](int d1, int s1, int d2, int s2) {
        int o[2] = {0, 0};
        o[d1] += s1;
        o[d2] += s2;
        return tile[v][(y + 2 + o[1]) * T + (x + 2 + o[0])];
      };
      double num = 0.0, den = 0.0;
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int d2 = 0; d2 < 2; ++d2) {
          const double tj = q(d2, 1, d, 1) - q(d2, 1, d, -1);
          const double th = q(d2, -1, d, 1) - q(d2, -1, d, -1);
          const double diff = tj - th;
          num = num + diff * diff;
        }
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const double q0 = q(d, 0, d, 0);
        const double tt = fabs(q(d, 2, d, 0) - q0) + fabs(q0 - q(d, -2, d, 0));
#pragma unroll
        for (int d2 = 0; d2 < 2; ++d2) {
          const double aj = fabs(q(d2, 1, d, 1)) + fabs(q(d2, 1, d, -1));
          const double ah = fabs(q(d2, -1, d, 1)) + fabs(q(d2, -1, d, -1));
          const double vv = tt + wavefilter * (aj + ah);
          den = den + vv * vv;
        }
      }
      error = error + wts[t0 + v] * sqrt(num / (den > 1.0e-6 ? den : 1.0e-6));
    }
    __syncthreads();  // the tile is rewritten for the next variables
  }
  const int any_r = __syncthreads_or(error >= thr ? 1 : 0);
  const int all_c = __syncthreads_and((error < thr && error <= cthr) ? 1 : 0);
  if (threadIdx.x == 0) {
    flags[b] = (unsigned char)((any_r ? 1 : 0) | (all_c ? 2 : 0));
    // the small-value flag of the stages before, so that one copy brings
    // both to the host
    if (b == 0) *crash_copy = *crash;
  }
}

// In 3-D the block with its stencil rim ((bx+4)^2 (bnz+4) cells, at most
// kLohner3Tile) is staged in LDS one variable at a time; a lane keeps the
// error sums of its (at most kLohner3Cells) cells in registers.
constexpr int kLohner3Tile = 8000, kLohner3Cells = 16;
__global__ void __launch_bounds__(NTHREADS) lohner3_kernel(SweepArgs a, int nw, int nwt, const int* __restrict__ iws,
                                                           const double* __restrict__ wts, double wavefilter,
                                                           double thr, double cthr, unsigned char* __restrict__ flags,
                                                           const int* __restrict__ crash, int* __restrict__ crash_copy) {
  __shared__ double tile[kLohner3Tile];
  const int b = blockIdx.x;
  const int n0 = a.bx, n2 = a.bnz;
  const int T0 = n0 + 4, T2 = n2 + 4;
  const int cells = n0 * n0 * n2, tcells = T0 * T0 * T2;
  double err[kLohner3Cells];
#pragma unroll
  for (int j = 0; j < kLohner3Cells; ++j) err[j] = 0.0;
  for (int t = 0; t < nwt; ++t) {
    const int iw = iws[t];
    for (int c = threadIdx.x; c < tcells; c += NTHREADS) {
      const int X = c % T0 - 2, Y = (c / T0) % T0 - 2, Z = c / (T0 * T0) - 2;
      const int o0 = X < 0 ? -X : (X >= n0 ? X - n0 + 1 : 0);
      const int o1 = Y < 0 ? -Y : (Y >= n0 ? Y - n0 + 1 : 0);
      const int o2 = Z < 0 ? -Z : (Z >= n2 ? Z - n2 + 1 : 0);
      const int nout = (o0 > 0) + (o1 > 0) + (o2 > 0);
      // the stencil reaches two cells out along an axis and one diagonally
      if (nout <= 1 || (nout == 2 && o0 <= 1 && o1 <= 1 && o2 <= 1)) tile[c] = cell_value<3>(a, nw, b, X, Y, Z, iw);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLohner3Cells; ++j) {
      const int c = threadIdx.x + j * NTHREADS;
      if (c >= cells) break;
      const int X[3] = {c % n0, (c / n0) % n0, c / (n0 * n0)};
      // q at p + s1 e_d1 + s2 e_d2, as the general path's q()
      auto q = [&](int d1, int s1, int d2, int s2) {
        int p[3] = {X[0] + 2, X[1] + 2, X[2] + 2};
        p[d1] += s1;
        p[d2] += s2;
        return tile[(p[2] * T0 + p[1]) * T0 + p[0]];
      };
      double num = 0.0, den = 0.0;
      for (int d = 0; d < 3; ++d)
        for (int d2 = 0; d2 < 3; ++d2) {
          const double tj = q(d2, 1, d, 1) - q(d2, 1, d, -1);
          const double th = q(d2, -1, d, 1) - q(d2, -1, d, -1);
          const double diff = tj - th;
          num = num + diff * diff;
        }
      for (int d = 0; d < 3; ++d) {
        const double q0 = q(d, 0, d, 0);
        const double tt = fabs(q(d, 2, d, 0) - q0) + fabs(q0 - q(d, -2, d, 0));
        for (int d2 = 0; d2 < 3; ++d2) {
          const double aj = fabs(q(d2, 1, d, 1)) + fabs(q(d2, 1, d, -1));
          const double ah = fabs(q(d2, -1, d, 1)) + fabs(q(d2, -1, d, -1));
          const double v = tt + wavefilter * (aj + ah);
          den = den + v * v;
        }
      }
      err[j] = err[j] + wts[t] * sqrt(num / (den > 1.0e-6 ? den : 1.0e-6));
    }
    __syncthreads();  // the tile is rewritten for the next variable
  }
  int any_r = 0, all_c = 1;
#pragma unroll
  for (int j = 0; j < kLohner3Cells; ++j) {
    if (threadIdx.x + j * NTHREADS >= cells) break;
    if (err[j] >= thr) any_r = 1;
    if (!(err[j] < thr && err[j] <= cthr)) all_c = 0;
  }
  any_r = __syncthreads_or(any_r);
  all_c = __syncthreads_and(all_c);
  if (threadIdx.x == 0) {
    flags[b] = (unsigned char)((any_r ? 1 : 0) | (all_c ? 2 : 0));
    if (b == 0) *crash_copy = *crash;
  }
}

template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) lohner_kernel(SweepArgs a, int nw, int nwt, const int* __restrict__ iws,
                                                          const double* __restrict__ wts, double wavefilter, double thr,
                                                          double cthr, unsigned char* __restrict__ flags) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  const int b = blockIdx.x;
  const int cells = BXr * BXr * a.bnz;
  int any_r = 0, all_c = 1;
  for (int c = threadIdx.x; c < cells; c += NTHREADS) {
    const int X[3] = {c % BXr, (c / BXr) % BXr, c / (BXr * BXr)};
    auto q = [&](int iw, int d1, int s1, int d2, int s2) {
      int p[3] = {X[0], X[1], X[2]};
      p[d1] += s1;
      p[d2] += s2;
      return cell_value<NDIM>(a, nw, b, p[0], p[1], p[2], iw);
    };
    double error = 0.0;
    for (int t = 0; t < nwt; ++t) {
      const int iw = iws[t];
      double num = 0.0, den = 0.0;
      for (int d = 0; d < NDIM; ++d)
        for (int d2 = 0; d2 < NDIM; ++d2) {
          const double tj = q(iw, d2, 1, d, 1) - q(iw, d2, 1, d, -1);
          const double th = q(iw, d2, -1, d, 1) - q(iw, d2, -1, d, -1);
          const double diff = tj - th;
          num = num + diff * diff;
        }
      for (int d = 0; d < NDIM; ++d) {
        const double q0 = q(iw, d, 0, d, 0);
        const double tt = fabs(q(iw, d, 2, d, 0) - q0) + fabs(q0 - q(iw, d, -2, d, 0));
        for (int d2 = 0; d2 < NDIM; ++d2) {
          const double aj = fabs(q(iw, d2, 1, d, 1)) + fabs(q(iw, d2, 1, d, -1));
          const double ah = fabs(q(iw, d2, -1, d, 1)) + fabs(q(iw, d2, -1, d, -1));
          const double v = tt + wavefilter * (aj + ah);
          den = den + v * v;
        }
      }
      error = error + wts[t] * sqrt(num / (den > 1.0e-6 ? den : 1.0e-6));
    }
    if (error >= thr) any_r = 1;
    if (!(error < thr && error <= cthr)) all_c = 0;
  }
  any_r = __syncthreads_or(any_r);
  all_c = __syncthreads_and(all_c);
  if (threadIdx.x == 0) flags[b] = (unsigned char)((any_r ? 1 : 0) | (all_c ? 2 : 0));
}

// Data of the new leaves after amr_coarsen_refine (src/amr/mod_coarsen_refine.t:
// 15-173), from the old leaves (a.src, old neighbour table).  plan[b*S ..]
// with S = 2 + 2^ndim: kind (0 copy, 1 coarsen, 2 refine), then the old
// block (copy, refine: the parent) or the 2^ndim old children in child order
// (coarsen), and last the child position sum_d (ic_d - 1) 2^d (refine).
//   coarsen: coarsen_grid (src/amr/mod_coarsen.t:12-79), the mean of the
//            2^ndim fine cells summed in array-element order;
//   refine:  prolong_2nd (src/amr/mod_refine.t:43-167) on conserved
//            variables: the coarse value plus limited slopes, the slope
//            of dimension d = sign(c) max(0, min(|c|, sign(c) l, sign(c) r))
//            with c = (l + r)/2, times eta = (offset - 1/2)/2.
template <int NDIM>
__global__ void __launch_bounds__(NTHREADS) remap_kernel(SweepArgs a, int nw, const int* __restrict__ plan,
                                                         double* __restrict__ out) {
  const int BXr = a.bx;  // block_nx(1) = block_nx(2)
  constexpr int S = 2 + (1 << NDIM);
  const int b = blockIdx.x;
  const int* p = plan + (size_t)b * S;
  const int kind = p[0];
  const int n[3] = {BXr, BXr, a.bnz};
  const size_t cells = (size_t)BXr * BXr * a.bnz;
  double* o = out + (size_t)b * nw * cells;
  for (int c = threadIdx.x; c < (int)cells; c += NTHREADS) {
    const int X[3] = {c % BXr, (c / BXr) % BXr, c / (BXr * BXr)};
    if (kind == 0) {
      const double* src = a.src + (size_t)p[1] * nw * cells;
      for (int iw = 0; iw < nw; ++iw) o[iw * cells + c] = src[iw * cells + c];
    } else if (kind == 1) {
      int f[3] = {0, 0, 0}, cidx = 0;
      for (int d = 0; d < NDIM; ++d) {
        const int ic = X[d] >= n[d] / 2 ? 2 : 1;
        cidx += (ic - 1) << d;
        f[d] = 2 * (X[d] - (ic - 1) * (n[d] / 2));
      }
      const double* wf = a.src + (size_t)p[1 + cidx] * nw * cells;
      for (int iw = 0; iw < nw; ++iw) {
        double sum = 0.0;
        bool first = true;
        for (int dk = 0; dk <= (NDIM > 2 ? 1 : 0); ++dk)
          for (int dj = 0; dj <= 1; ++dj)
            for (int di = 0; di <= 1; ++di) {
              const double v = wf[iw * cells + ((size_t)(f[2] + dk) * BXr + (f[1] + dj)) * BXr + f[0] + di];
              sum = first ? v : sum + v;
              first = false;
            }
        o[iw * cells + c] = sum * (1.0 / double(1 << NDIM));
      }
    } else {
      const int P = p[1], icode = p[S - 1];
      int cc[3] = {0, 0, 0}, off[3] = {0, 0, 0};
      for (int d = 0; d < NDIM; ++d) {
        const int ic = 1 + ((icode >> d) & 1);
        cc[d] = (ic - 1) * (n[d] / 2) + X[d] / 2;
        off[d] = X[d] % 2;
      }
      for (int iw = 0; iw < nw; ++iw) {
        const double c0 = cell_value<NDIM>(a, nw, P, cc[0], cc[1], cc[2], iw);
        double v = c0;
        for (int d = 0; d < NDIM; ++d) {
          int m[3] = {cc[0], cc[1], cc[2]}, pl[3] = {cc[0], cc[1], cc[2]};
          m[d] -= 1;
          pl[d] += 1;
          const double sl = c0 - cell_value<NDIM>(a, nw, P, m[0], m[1], m[2], iw);
          const double sr = cell_value<NDIM>(a, nw, P, pl[0], pl[1], pl[2], iw) - c0;
          const double sc = 0.5 * (sr + sl);
          const double sg = copysign(1.0, sc);
          double lim = fabs(sc);
          if (sg * sl < lim) lim = sg * sl;
          if (sg * sr < lim) lim = sg * sr;
          const double slope = sg * (lim > 0.0 ? lim : 0.0);
          v = v + slope * (0.5 * (double(off[d]) - 0.5));
        }
        o[iw * cells + c] = v;
      }
    }
  }
}
}  // namespace fv

namespace {

fvc::KernelKey make_key(const amrvac_fv_ctx* c, int riem) {
  fvc::KernelKey k{};
  k.ndim = c->cfg.ndim;
  k.physics = c->cfg.physics;
  k.ndir = c->cfg.ndir;
  // constrained transport and Linde's cleaning: the stage kernels are those
  // of ideal MHD (Linde's source follows in linde_kernel)
  k.divb = c->cfg.physics == 2 ? ((c->cfg.divb == 8 || c->cfg.divb == 4) ? 0 : c->cfg.divb) : 0;
  k.mixed = c->cfg.precision == AMRVAC_FV_PREC_MIXED ? 1 : 0;
  k.riem = riem;
  k.weno = c->cfg.limiter == 16 ? 1 : 0;
  k.limiter = c->cfg.limiter;
  return k;
}

hipError_t dispatch_sweep(const amrvac_fv_ctx* c, const fvc::KernelKey& k, const fvc::SweepArgs& a, int grid) {
  const bool exact = c->cfg.precision == 2;
  if (k.ndim == 2 && k.physics == 1)
    return exact ? fvx::launch_sweep_2d_hd(k, a, c->fp, grid, c->stream)
                 : fv::launch_sweep_2d_hd(k, a, c->fp, grid, c->stream);
  if (k.ndim == 2 && k.physics == 2)
    return exact ? fvx::launch_sweep_2d_mhd(k, a, c->fp, grid, c->stream)
                 : fv::launch_sweep_2d_mhd(k, a, c->fp, grid, c->stream);
  if (k.ndim == 3 && k.physics == 1)
    return exact ? fvx::launch_sweep_3d_hd(k, a, c->fp, grid, c->stream)
                 : fv::launch_sweep_3d_hd(k, a, c->fp, grid, c->stream);
  if (k.ndim == 3 && k.physics == 2)
    return exact ? fvx::launch_sweep_3d_mhd(k, a, c->fp, grid, c->stream)
                 : fv::launch_sweep_3d_mhd(k, a, c->fp, grid, c->stream);
  return hipErrorInvalidValue;
}

template <template <class> class F, class... Args>
hipError_t with_phys(const amrvac_fv_ctx* c, Args... args) {
  using namespace fv;
  // constrained transport uses the ideal-MHD physics without a source term
  // and so does Linde's cleaning, whose source is a kernel of its own
  const int nd = c->cfg.ndim, ph = c->cfg.physics, nr = c->cfg.ndir;
  const int db = (c->cfg.divb == 8 || c->cfg.divb == 4) ? 0 : c->cfg.divb;
  if (nd == 2 && ph == 1 && nr == 2) return F<Phys<2, 2, false, DIVB_NONE>>::run(args...);
  if (nd == 2 && ph == 2 && nr == 3 && db == DIVB_GLM) return F<Phys<2, 3, true, DIVB_GLM>>::run(args...);
  if (nd == 2 && ph == 2 && nr == 3 && db == DIVB_NONE) return F<Phys<2, 3, true, DIVB_NONE>>::run(args...);
  if (nd == 2 && ph == 2 && nr == 2 && db == DIVB_GLM) return F<Phys<2, 2, true, DIVB_GLM>>::run(args...);
  if (nd == 2 && ph == 2 && nr == 2 && db == DIVB_NONE) return F<Phys<2, 2, true, DIVB_NONE>>::run(args...);
  if (nd == 3 && ph == 1 && nr == 3) return F<Phys<3, 3, false, DIVB_NONE>>::run(args...);
  if (nd == 3 && ph == 2 && nr == 3 && db == DIVB_GLM) return F<Phys<3, 3, true, DIVB_GLM>>::run(args...);
  if (nd == 3 && ph == 2 && nr == 3 && db == DIVB_NONE) return F<Phys<3, 3, true, DIVB_NONE>>::run(args...);
  return hipErrorInvalidValue;
}

template <class P>
struct LindeOp {
  static hipError_t run(const fvc::SweepArgs* a, const fvc::FVParams* fp, const int* ntype, int nnb, double* dst,
                        hipStream_t s) {
    return fv::launch_linde<P>(*a, *fp, ntype, nnb, dst, s);
  }
};

template <class P>
struct GetdtOp {
  static hipError_t run(const double* w, int nb, int bx, int bnz, const fvc::FVParams* fp, const int* blevel,
                        double* out, int* crash, hipStream_t s) {
    return fv::launch_getdt<P>(w, nb, bx, bnz, *fp, blevel, out, crash, s);
  }
};

bool check_slot(int s) { return s >= 0 && s <= 3; }

}  // namespace

// grow a device buffer of the context to hold n elements
template <class T>
static bool ensure_buf(T** buf, size_t* cap, size_t n) {
  if (*buf && *cap >= n) return true;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc(buf, n * sizeof(T)) != hipSuccess) return false;
  *cap = n;
  return true;
}

constexpr size_t kPinCrash = 24, kPinFlags = 32;

static bool ensure_pin(amrvac_fv_ctx* c, size_t bytes) {
  if (c->h_pin && c->pin_cap >= bytes) return true;
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  c->h_pin = nullptr;
  c->pin_cap = 0;
  const size_t n = bytes + bytes / 4 + 64;
  if (hipHostMalloc((void**)&c->h_pin, n) != hipSuccess) {
    c->h_pin = nullptr;
    return false;
  }
  c->pin_cap = n;
  return true;
}

// the small-value flag, read back through the pinned buffer
static hipError_t read_crash(amrvac_fv_ctx* c, int* out) {
  int* h = reinterpret_cast<int*>(c->h_pin + kPinCrash);
  hipError_t e = hipMemcpyAsync(h, c->d_crash, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) *out = *h;
  return e;
}

// a table of the level into its buffer (grown when it does not fit)
template <class T>
static bool upload_vec(amrvac_fv_ctx* c, T** dst, size_t* cap, const std::vector<T>& v) {
  if (v.empty()) return true;
  if (!ensure_buf(dst, cap, v.size())) return false;
  return hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream) == hipSuccess;
}

extern "C" {

const char* amrvac_fv_last_error(void) { return g_err.c_str(); }

int amrvac_fv_init(const amrvac_fv_config* cfg, amrvac_fv_ctx** out) {
  if (!cfg || !out) return fail(AMRVAC_FV_EINVAL, "null argument");
  const amrvac_fv_config& c = *cfg;
  if (c.ndim < 2 || c.ndim > 3) return fail(AMRVAC_FV_EINVAL, "ndim must be 2 or 3");
  if (c.ndir < c.ndim || c.ndir > 3) return fail(AMRVAC_FV_EINVAL, "need ndim <= ndir <= 3");
  if (c.physics != 1 && c.physics != 2) return fail(AMRVAC_FV_EINVAL, "physics must be 1 (hd) or 2 (mhd)");
  if (c.physics == 1 && c.ndir != c.ndim) return fail(AMRVAC_FV_EINVAL, "hd: ndir must equal ndim");
  if (c.physics == 2 && !(c.ndir == 3 || (c.ndir == 2 && c.ndim == 2)))
    return fail(AMRVAC_FV_EINVAL, "mhd: ndir must be 3 (or 2 in 2-D)");
  if (c.physics == 2 && c.divb != 0 && c.divb != 1 && c.divb != 4 && c.divb != 8)
    return fail(AMRVAC_FV_EINVAL, "typedivbfix must be none (0), glm (1), linde (4) or ct (8)");
  if (c.physics == 2 && c.divb == 8) {
    if (c.ndir != 3) return fail(AMRVAC_FV_EINVAL, "constrained transport needs ndir = 3");
    if (c.block_nx[0] != 16 || (c.ndim == 3 && c.block_nx[2] != 16))
      return fail(AMRVAC_FV_EINVAL, "constrained transport needs block_nx = 16");
    if (c.type_ct < 1 || c.type_ct > 3) return fail(AMRVAC_FV_EINVAL, "type_ct must be 1, 2 or 3");
    if (c.limiter == 16) return fail(AMRVAC_FV_EINVAL, "constrained transport takes a TVD limiter");
  }
  if (c.block_nx[0] != c.block_nx[1])
    return fail(AMRVAC_FV_EINVAL, "block_nx(1) and block_nx(2) must be equal");
  if (c.block_nx[0] != 16) {
    if (c.block_nx[0] != 8 && c.block_nx[0] != 10)
      return fail(AMRVAC_FV_EINVAL, "block_nx(1) = block_nx(2) must be 16, or 8 or 10 in 3-D");
    if (c.ndim != 3) return fail(AMRVAC_FV_EINVAL, "block_nx(1) = block_nx(2) must be 16 in 2-D");
    if (c.precision == AMRVAC_FV_PREC_MIXED)
      return fail(AMRVAC_FV_EINVAL, "the mixed form needs block_nx(1) = block_nx(2) = 16");
  }
  if (c.ndim == 3 && c.block_nx[2] < 4) return fail(AMRVAC_FV_EINVAL, "block_nx(3) must be at least 4");
  if (c.physics == 2 && c.divb == 4 &&
      (c.block_nx[0] + 2) * (c.block_nx[0] + 2) * (c.ndim == 3 ? c.block_nx[2] + 2 : 1) > fv::kLindeLds)
    return fail(AMRVAC_FV_EINVAL, "Linde cleaning: block_nx(3) too large for the div B tile (at most 22 with "
                                  "16^2 cross-sections)");
  const bool weno = c.limiter == 16;
  if (!weno && !(c.limiter >= 1 && c.limiter <= 9))
    return fail(AMRVAC_FV_EINVAL, "unsupported limiter code");
  if (c.nghost != (weno ? 3 : 2))
    return fail(AMRVAC_FV_EINVAL, weno ? "weno5 needs nghostcells = 3" : "nghostcells must be 2");
  if (c.boundspeed < 1 || c.boundspeed > 3) return fail(AMRVAC_FV_EINVAL, "boundspeed must be 1, 2 or 3");
  if (c.precision < 0 || c.precision > 2) return fail(AMRVAC_FV_EINVAL, "precision must be 0, 1 or 2");
  if (c.gamma <= 0.0 || c.gamma == 1.0) return fail(AMRVAC_FV_EINVAL, "gamma must be > 0 and != 1");
  if (c.fix_small_values != 0 && c.fix_small_values != 1)
    return fail(AMRVAC_FV_EINVAL, "fix_small_values must be 0 or 1");
  if (c.fix_small_values && c.physics == 2 && c.divb == 8)
    return fail(AMRVAC_FV_EINVAL, "fix_small_values is not provided with constrained transport");
  amrvac_fv_ctx* x = new amrvac_fv_ctx();
  x->cfg = c;
  x->nw = c.physics == 1 ? c.ndir + 2 : 2 * c.ndir + 2 + (c.divb == 1 ? 1 : 0);
  x->nws = (c.physics == 2 && c.divb == 8) ? c.ndim : 0;
  x->nnb = c.ndim == 2 ? 9 : 27;
  x->bnz = c.ndim == 3 ? c.block_nx[2] : 1;
  x->bx = c.block_nx[0];
  x->cells = (size_t)x->bx * x->bx * x->bnz;
  for (int d = 0; d < 3; ++d) x->nxg[d] = d < c.ndim ? c.block_nx[d] + 2 * c.nghost : 1;
  x->cells_g = (size_t)x->nxg[0] * x->nxg[1] * x->nxg[2];
  hipError_t e = hipStreamCreate(&x->stream);
  if (e != hipSuccess) {
    delete x;
    return fail(AMRVAC_FV_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  x->own_stream = true;
  const int rc = scratch_address(&x->scratch);
  if (rc != AMRVAC_FV_OK) {
    (void)hipStreamDestroy(x->stream);
    delete x;
    return rc;
  }
  *out = x;
  return AMRVAC_FV_OK;
}


static void free_level(amrvac_fv_ctx* c) {
  for (int s = 0; s < 4; ++s) { if (c->slot[s]) (void)hipFree(c->slot[s]); c->slot[s] = nullptr; }
  for (int s = 0; s < 4; ++s) { if (c->wsl[s]) (void)hipFree(c->wsl[s]); c->wsl[s] = nullptr; }
  if (c->d_ctf) (void)hipFree(c->d_ctf);
  if (c->d_wsn) (void)hipFree(c->d_wsn);
  c->d_ctf = nullptr;
  c->d_wsn = nullptr;
  if (c->d_nbr) (void)hipFree(c->d_nbr);
  if (c->d_crash) (void)hipFree(c->d_crash);
  if (c->d_red) (void)hipFree(c->d_red);
  if (c->d_red2) (void)hipFree(c->d_red2);
  if (c->d_dtp) (void)hipFree(c->d_dtp);
  c->d_dtp = nullptr;
  c->dt_valid = false;
  if (c->d_stage) (void)hipFree(c->d_stage);
  c->d_nbr = nullptr; c->d_crash = nullptr; c->d_red = nullptr; c->d_red2 = nullptr; c->d_stage = nullptr;
  void* amr_bufs[] = {c->d_level, c->d_ntype, c->d_nchild, c->d_rtask, c->d_ptask, c->d_ptask2,
                      c->d_ftask, c->d_fregoff, c->d_freg, c->d_xlo, c->d_adjG};
  for (void* p : amr_bufs)
    if (p) (void)hipFree(p);
  c->d_level = nullptr; c->d_ntype = nullptr; c->d_nchild = nullptr; c->d_rtask = nullptr;
  c->d_ptask = nullptr; c->d_ptask2 = nullptr; c->d_ftask = nullptr; c->d_fregoff = nullptr;
  c->d_freg = nullptr; c->d_xlo = nullptr; c->d_adjG = nullptr;
  c->slot_cap = c->nbr_cap = c->red_cap = c->dtp_cap = c->stage_cap = 0;
  c->level_cap = c->ntype_cap = c->nchild_cap = c->rtask_cap = c->ptask_cap = c->ptask2_cap = 0;
  c->ftask_cap = c->fregoff_cap = c->freg_cap = c->xlo_cap = 0;
  c->amr = false;
  c->nv = 0;
  c->nrt = c->npt = c->nft = c->nreg = 0;
  c->h_level.clear();
}

// an AMR level replaced by the next one (regrid): the old level's state is
// dropped, its buffers are kept for alloc_level to reuse where they fit
static void reset_level(amrvac_fv_ctx* c) {
  if (c->d_adjG) (void)hipFree(c->d_adjG);
  c->d_adjG = nullptr;
  c->dt_valid = false;
  c->amr = false;
  c->nv = 0;
  c->nrt = c->npt = c->nft = c->nreg = 0;
  c->h_level.clear();
}

int amrvac_fv_finalize(amrvac_fv_ctx* c) {
  if (!c) return AMRVAC_FV_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_level(c);
  void* regrid_bufs[] = {c->d_remap, c->d_plan, c->d_flags, c->d_iws, c->d_wts};
  for (void* p : regrid_bufs)
    if (p) (void)hipFree(p);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  for (auto& ev : c->events) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AMRVAC_FV_OK;
}

int amrvac_fv_set_stream(amrvac_fv_ctx* c, void* stream) {
  if (!c) return fail(AMRVAC_FV_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (stream) {
    if (c->own_stream) HIPCHK(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else if (!c->own_stream) {
    HIPCHK(hipStreamCreate(&c->stream));
    c->own_stream = true;
  }
  return AMRVAC_FV_OK;
}

static int alloc_level(amrvac_fv_ctx* c, int nblocks, int capacity, const double* dx, const int* neighbor,
                       const int* bc);

int amrvac_fv_set_mesh(amrvac_fv_ctx* c, int nblocks, const double* dx, const int* neighbor,
                       const double* xlohi, const int* bc) {
  if (!c || !dx || !neighbor || !bc) return fail(AMRVAC_FV_EINVAL, "null argument");
  (void)xlohi;
  if (nblocks < 1) return fail(AMRVAC_FV_EINVAL, "nblocks must be positive");
  const int nd = c->cfg.ndim;
  for (int i = 0; i < nblocks * c->nnb; ++i)
    if (neighbor[i] >= nblocks) return fail(AMRVAC_FV_EINVAL, "neighbour index out of range");
  for (int b = 0; b < nblocks; ++b)
    if (neighbor[b * c->nnb + (c->nnb - 1) / 2] != b)
      return fail(AMRVAC_FV_EINVAL, "neighbour table centre entry must be the block itself");
  for (int s = 0; s < 2 * nd; ++s)
    if (bc[s] != 2 && bc[s] != 3 && bc[s] != 4 && bc[s] != 5)
      return fail(AMRVAC_FV_EINVAL, "boundary type must be cont, symm, asymm or periodic");
  for (int d = 0; d < nd; ++d)
    if (!(dx[d] > 0.0)) return fail(AMRVAC_FV_EINVAL, "dx must be positive");
  if (c->nws > 0)
    for (int i = 0; i < nblocks * c->nnb; ++i)
      if (neighbor[i] < 0) return fail(AMRVAC_FV_EINVAL, "constrained transport needs a periodic level");
  HIPCHK(hipStreamSynchronize(c->stream));
  free_level(c);
  return alloc_level(c, nblocks, nblocks, dx, neighbor, bc);
}

// state slots with room for `capacity` blocks (leaves + virtual ghost
// blocks), work buffers, the sweep's neighbour table and the scheme
// constants of the level-1 cell size dx
static int alloc_level(amrvac_fv_ctx* c, int nblocks, int capacity, const double* dx, const int* neighbor,
                       const int* bc) {
  const int nd = c->cfg.ndim;
  c->nb = nblocks;
  for (int d = 0; d < 3; ++d) c->dx[d] = d < nd ? dx[d] : 1.0;
  for (int s = 0; s < 6; ++s) c->bc[s] = s < 2 * nd ? bc[s] : 5;
  const size_t n = (size_t)capacity * c->nw * c->cells;
  if (c->slot_cap < n) {
    for (int s = 0; s < 4; ++s) {
      if (c->slot[s]) (void)hipFree(c->slot[s]);
      c->slot[s] = nullptr;
    }
    c->slot_cap = 0;
    for (int s = 0; s < 4; ++s)
      if (hipMalloc(&c->slot[s], n * sizeof(double)) != hipSuccess) {
        free_level(c);
        return fail(AMRVAC_FV_ENOMEM, "state slot allocation failed");
      }
    c->slot_cap = n;
  }
  for (int s = 0; s < 4; ++s) HIPCHK(hipMemsetAsync(c->slot[s], 0, n * sizeof(double), c->stream));
  if (c->nws > 0) {
    const size_t nf = (size_t)capacity * c->nws * c->cells;
    for (int s = 0; s < 4; ++s) {
      if (hipMalloc(&c->wsl[s], nf * sizeof(double)) != hipSuccess) {
        free_level(c);
        return fail(AMRVAC_FV_ENOMEM, "face field allocation failed");
      }
      HIPCHK(hipMemsetAsync(c->wsl[s], 0, nf * sizeof(double), c->stream));
    }
    const size_t nfb = nd == 3 ? fv::ct3_scratch_doubles(nblocks) : fv::ct_face_buffer_doubles(nblocks);
    if (hipMalloc(&c->d_ctf, nfb * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->d_wsn, nf * sizeof(double)) != hipSuccess) {
      free_level(c);
      return fail(AMRVAC_FV_ENOMEM, "constrained-transport buffers allocation failed");
    }
  }
  if (!ensure_buf(&c->d_nbr, &c->nbr_cap, (size_t)nblocks * c->nnb) ||
      (!c->d_crash && hipMalloc(&c->d_crash, sizeof(int)) != hipSuccess) ||
      !ensure_buf(&c->d_red, &c->red_cap, (size_t)nblocks * 2 * (c->nw > 1 ? c->nw : 1)) ||
      (!c->d_red2 && hipMalloc(&c->d_red2, sizeof(double) * 3) != hipSuccess) || !ensure_pin(c, kPinFlags + nblocks) ||
      !ensure_buf(&c->d_dtp, &c->dtp_cap, (size_t)nblocks * 2) ||
      !ensure_buf(&c->d_stage, &c->stage_cap, (size_t)nblocks * c->nw * c->cells_g)) {
    free_level(c);
    return fail(AMRVAC_FV_ENOMEM, "work buffer allocation failed");
  }
  HIPCHK(hipMemcpyAsync(c->d_nbr, neighbor, sizeof(int) * nblocks * c->nnb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->d_crash, 0, sizeof(int), c->stream));
  c->crash_seen = 0;
  c->crash_fresh = false;
  // scheme constants that depend on the mesh
  fvc::FVParams& fp = c->fp;
  const amrvac_fv_config& g = c->cfg;
  fp.gamma = g.gamma;
  fp.gamma_1 = g.gamma - 1.0;
  fp.inv_gamma_1 = 1.0 / (g.gamma - 1.0);
  fp.glm_alpha = g.glm_alpha;
  fp.glm_extended = g.glm_extended;
  fp.divbdiff = g.divbdiff;
  fp.limiter = g.limiter;
  fp.boundspeed = g.boundspeed;
  fp.tvdlfeps = g.tvdlfeps;
  fp.small_pressure = g.small_pressure;
  fp.small_density = g.small_density;
  fp.check_small = g.check_small_values;
  fp.fix_small = g.fix_small_values;
  fp.invds2 = 0.0;
  fp.mindx = c->dx[0];
  for (int d = 0; d < 3; ++d) {
    fp.dx[d] = c->dx[d];
    fp.dxinv[d] = 1.0 / c->dx[d];
    const double r = g.cada3_radius * c->dx[d];
    fp.rdelinv[d] = 1.0 / (r * r);
    if (d < nd) {
      fp.invds2 = fp.invds2 + 1.0 / (c->dx[d] * c->dx[d]);
      if (c->dx[d] < fp.mindx) fp.mindx = c->dx[d];
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRVAC_FV_OK;
}

int amrvac_fv_upload(amrvac_fv_ctx* c, int slot, const double* w) {
  if (!c || !w || !check_slot(slot)) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (slot == 0) c->dt_valid = false;
  const int ng = c->cfg.nghost, nd = c->cfg.ndim;
  const int nwt = c->nw + c->nws;  // with constrained transport the face field follows the cell variables
  const size_t n = (size_t)c->nb * nwt * c->cells;
  c->h_stage.resize(n);
  for (int b = 0; b < c->nb; ++b)
    for (int iw = 0; iw < nwt; ++iw) {
      const double* src = w + ((size_t)b * nwt + iw) * c->cells_g;
      // cell variables first (all blocks), then the face components
      double* dst = iw < c->nw ? c->h_stage.data() + ((size_t)b * c->nw + iw) * c->cells
                               : c->h_stage.data() + (size_t)c->nb * c->nw * c->cells +
                                     ((size_t)b * c->nws + (iw - c->nw)) * c->cells;
      for (int k = 0; k < c->bnz; ++k)
        for (int j = 0; j < c->bx; ++j) {
          const int kg = nd > 2 ? k + ng : 0;
          const double* row = src + ((size_t)kg * c->nxg[1] + (j + ng)) * c->nxg[0] + ng;
          std::memcpy(dst + ((size_t)k * c->bx + j) * c->bx, row, c->bx * sizeof(double));
        }
    }
  const size_t nc = (size_t)c->nb * c->nw * c->cells;
  HIPCHK(hipMemcpyAsync(c->slot[slot], c->h_stage.data(), nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->nws > 0)
    HIPCHK(hipMemcpyAsync(c->wsl[slot], c->h_stage.data() + nc, (n - nc) * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRVAC_FV_OK;
}

static fvc::SweepArgs base_args(const amrvac_fv_ctx* c) {
  fvc::SweepArgs a{};
  a.nbr = c->d_nbr;
  a.nblocks = c->nb;
  a.bx = c->bx;
  a.bnz = c->bnz;
  a.per_xcd = (c->nb + 7) / 8;
  for (int s = 0; s < 6; ++s) a.bc[s] = c->bc[s];
  a.crash = c->d_crash;
  return a;
}


}  // extern "C"

// ghost cells across level jumps of a slot: restriction into the coarse
// blocks' virtual blocks, then prolongation into the fine blocks' (which
// reads the former through the coarse block's neighbour table)
static int getbc_amr(amrvac_fv_ctx* c, int slot) {
  if (!c->amr) return AMRVAC_FV_OK;
  const int nd = c->cfg.ndim, ng = c->cfg.nghost;
  if (c->nrt) {
    if (nd == 2)
      hipLaunchKernelGGL((fv::restrict_kernel<2>), dim3(c->nrt), dim3(fvc::NTHREADS), 0, c->stream, c->slot[slot],
                         c->nw, c->bx, c->bnz, ng, c->nnb, c->d_rtask, c->d_nchild);
    else
      hipLaunchKernelGGL((fv::restrict_kernel<3>), dim3(c->nrt), dim3(fvc::NTHREADS), 0, c->stream, c->slot[slot],
                         c->nw, c->bx, c->bnz, ng, c->nnb, c->d_rtask, c->d_nchild);
    HIPCHK(hipGetLastError());
  }
  if (c->npt) {
    fvc::SweepArgs a = base_args(c);
    a.src = c->slot[slot];
    if (nd == 2)
      hipLaunchKernelGGL((fv::prolong_kernel<2>), dim3(c->npt), dim3(fvc::NTHREADS), 0, c->stream, a, c->slot[slot],
                         c->nw, ng, c->nnb, c->d_ptask, c->d_ptask2, c->d_ntype, c->d_level, c->d_xlo, c->fp);
    else
      hipLaunchKernelGGL((fv::prolong_kernel<3>), dim3(c->npt), dim3(fvc::NTHREADS), 0, c->stream, a, c->slot[slot],
                         c->nw, ng, c->nnb, c->d_ptask, c->d_ptask2, c->d_ntype, c->d_level, c->d_xlo, c->fp);
    HIPCHK(hipGetLastError());
  }
  return AMRVAC_FV_OK;
}

static int fix_conserve_amr(amrvac_fv_ctx* c, int slot) {
  if (!c->amr || !c->fix_conserve || !c->nft) return AMRVAC_FV_OK;
  if (c->cfg.ndim == 2)
    hipLaunchKernelGGL((fv::fixcons_kernel<2>), dim3(c->nft), dim3(fvc::NTHREADS), 0, c->stream, c->slot[slot],
                       c->nw, c->bx, c->bnz, c->nnb, c->d_ftask, c->d_ntype, c->d_nchild, c->d_fregoff, c->d_freg,
                       c->nface);
  else
    hipLaunchKernelGGL((fv::fixcons_kernel<3>), dim3(c->nft), dim3(fvc::NTHREADS), 0, c->stream, c->slot[slot],
                       c->nw, c->bx, c->bnz, c->nnb, c->d_ftask, c->d_ntype, c->d_nchild, c->d_fregoff, c->d_freg,
                       c->nface);
  HIPCHK(hipGetLastError());
  return AMRVAC_FV_OK;
}

// reduction of the per-block CFL maxima part[2*nb]: their max and the
// small-value flag into fin[3], copied to the host; dt and cmax from them
static int getdt_reduce(amrvac_fv_ctx* c, double courantpar, double* dtmin_out, double* cmax_out, double* part,
                        double* fin);

// dt of state u: per-block maxima into part[2*nb], then getdt_reduce
static int getdt_core(amrvac_fv_ctx* c, const double* u, double courantpar, double* dtmin_out, double* cmax_out,
                      double* part, double* fin) {
  hipError_t e = with_phys<GetdtOp>(c, u, c->nb, c->bx, c->bnz, (const fvc::FVParams*)&c->fp, (const int*)c->d_level,
                                    part, c->d_crash, c->stream);
  if (e != hipSuccess) return fail(AMRVAC_FV_EHIP, std::string("getdt launch: ") + hipGetErrorString(e));
  return getdt_reduce(c, courantpar, dtmin_out, cmax_out, part, fin);
}

static int getdt_reduce(amrvac_fv_ctx* c, double courantpar, double* dtmin_out, double* cmax_out, double* part,
                        double* fin) {
  hipLaunchKernelGGL(fv::reduce_max2_kernel, dim3(1), dim3(fvc::NTHREADS), 0, c->stream, part, c->nb, fin,
                     (const int*)c->d_crash);
  HIPCHK(hipGetLastError());
  const double* r = reinterpret_cast<const double*>(c->h_pin);
  HIPCHK(hipMemcpyAsync(c->h_pin, fin, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the small-value flag as of this reduction; valid until the next launch
  // that may set it
  c->crash_seen = r[2] != 0.0 ? 1 : 0;
  c->crash_fresh = true;
  // min over blocks of courantpar/maxtot equals courantpar/max(maxtot): the
  // correctly rounded quotient is monotone in the divisor
  const double dt = r[0] > kSmall ? courantpar / r[0] : kBig;
  if (dtmin_out) *dtmin_out = dt < kBig ? dt : kBig;
  if (cmax_out) *cmax_out = r[1];
  return AMRVAC_FV_OK;
}

static int method_to_riem(int method, int& riem) {
  switch (method) {
    case 1: riem = fv::RIEM_HLL; return 0;
    case 2: riem = fv::RIEM_HLLC; return 0;
    case 3: riem = fv::RIEM_HLLD; return 0;
    case 5: riem = fv::RIEM_TVDLF; return 0;
    case 8: riem = fv::RIEM_HANCOCK; return 0;
    default: return -1;
  }
}

// One RK stage, dst = init + L(src) with init per comb_mode (0: dst,
// 1: base, 2: comb_a*base + comb_b*src, 3: src); the AMR hooks
// (refluxing, ghost cells of dst_slot) run when dst_slot >= 0.
static int stage_core(amrvac_fv_ctx* c, int method, double dtfactor, double dt, double cmax_global,
                      const double* src, double* dst, const double* base, int comb_mode, double comb_a,
                      double comb_b, int dst_slot, int* crash_out) {
  int riem;
  if (method_to_riem(method, riem) != 0) return fail(AMRVAC_FV_EINVAL, "unsupported flux scheme");
  if (riem == fv::RIEM_HLLD && (c->cfg.physics != 2 || c->cfg.ndir != 3))
    return fail(AMRVAC_FV_EINVAL, "hlld is provided for mhd with ndir = 3");
  if (riem == fv::RIEM_HANCOCK && c->cfg.limiter == 16)
    return fail(AMRVAC_FV_EINVAL, "the Hancock predictor takes a TVD limiter");
  if (c->cfg.physics == 2 && c->cfg.divb == 1 && !(cmax_global > 0.0))
    return fail(AMRVAC_FV_EINVAL, "GLM needs cmax_global > 0");
  fvc::SweepArgs a = base_args(c);
  a.src = src;
  a.dst = dst;
  a.base = base;
  a.comb_mode = comb_mode;
  a.comb_a = comb_a;
  a.comb_b = comb_b;
  const double qdt = dtfactor * dt;
  a.qdt = qdt;
  for (int d = 0; d < 3; ++d) a.dq[d] = -qdt / c->dx[d];
  a.ch = cmax_global;
  // add_source_glm psi decay factor, src/mhd/mod_mhd_phys.t:5491-5495
  a.glm_decay = std::exp(-qdt * cmax_global * c->cfg.glm_alpha / c->fp.mindx);
  if (c->amr) {
    a.blevel = c->d_level;
    double mindx = c->fp.mindx;
    for (int l = 0; l < fvc::MAXLEV; ++l) {
      a.glm_decay_lev[l] = std::exp(-qdt * cmax_global * c->cfg.glm_alpha / mindx);
      mindx = mindx * 0.5;
    }
    if (c->fix_conserve && c->nreg > 0 && riem != fv::RIEM_HANCOCK) {
      a.freg = c->d_freg;
      a.fregoff = c->d_fregoff;
      a.nface = c->nface;
    }
  }
  const fvc::KernelKey k = make_key(c, riem);
  c->crash_fresh = false;
  // ps written on a uniform level by the mixed-form kernels: the stage also
  // leaves the CFL partials of its output for the next getdt (the dt
  // reduction of setdt, src/mod_dt.t:160-287, without a pass over ps).
  // The FP64 kernels do not compute them (the extra work in their z loop
  // costs more than the separate pass), nor do the AMR-capable kernels of
  // 8^2 / 10^2 cross-sections.
  const bool linde = c->cfg.physics == 2 && c->cfg.divb == 4;
  a.linde = linde ? 1 : 0;
  const bool fuse_dt = dst_slot == 0 && !c->amr && c->bx == 16 && c->nws == 0 && !linde &&
                       c->cfg.precision == AMRVAC_FV_PREC_MIXED && !c->slot0_exposed;
  if (fuse_dt) a.dt_part = c->d_dtp;
  if (dst_slot == 0 || dst_slot < 0) c->dt_valid = false;
  if (crash_out) HIPCHK(hipMemsetAsync(c->d_crash, 0, sizeof(int), c->stream));
  const int grid = 8 * a.per_xcd;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    if (c->nev == c->events.size()) {
      hipEvent_t x0, x1;
      HIPCHK(hipEventCreate(&x0));
      HIPCHK(hipEventCreate(&x1));
      c->events.push_back({x0, x1});
    }
    e0 = c->events[c->nev].first;
    e1 = c->events[c->nev].second;
    ++c->nev;
    HIPCHK(hipEventRecord(e0, c->stream));
  }
  hipError_t e = dispatch_sweep(c, k, a, grid);
  if (e == hipErrorInvalidValue) return fail(AMRVAC_FV_EINVAL, "no stage kernel for this configuration");
  if (e != hipSuccess) return fail(AMRVAC_FV_EHIP, std::string("stage kernel: ") + hipGetErrorString(e));
  if (c->timing) HIPCHK(hipEventRecord(e1, c->stream));
  if (linde) {  // addsource2 -> add_source_linde, before refluxing
    e = with_phys<LindeOp>(c, (const fvc::SweepArgs*)&a, (const fvc::FVParams*)&c->fp,
                           (const int*)(c->amr ? c->d_ntype : nullptr), c->nnb, dst, c->stream);
    if (e != hipSuccess) return fail(AMRVAC_FV_EHIP, std::string("Linde source kernel: ") + hipGetErrorString(e));
  }
  if (fuse_dt) c->dt_valid = true;
  if (c->amr && dst_slot >= 0) {
    // refluxing, then the ghost cells of dst (advect1: fix_conserve, getbc)
    if (riem != fv::RIEM_HANCOCK) {
      const int rc = fix_conserve_amr(c, dst_slot);
      if (rc != AMRVAC_FV_OK) return rc;
    }
    const int rc = getbc_amr(c, dst_slot);
    if (rc != AMRVAC_FV_OK) return rc;
  }
  if (crash_out) HIPCHK(read_crash(c, crash_out));
  return AMRVAC_FV_OK;
}

// One constrained-transport stage between slots (advect1 with stagger_grid,
// uniform periodic 2-D or 3-D level): face fluxes, edge fields, face and cell
// updates; the new face field replaces psb's (an array swap).  ps's face field feeds
// the HLLD star states; when psb is ps it holds the stage's initial
// combination.
static int ct_stage(amrvac_fv_ctx* c, int method, double dtfactor, double dt, int psa, int psb, int comb_slot,
                    int mode, double comb_a, double comb_b, int* crash_out) {
  int riem;
  if (method_to_riem(method, riem) != 0) return fail(AMRVAC_FV_EINVAL, "unsupported flux scheme");
  if (riem == fv::RIEM_HANCOCK) {
    // hancock (src/mod_finite_volume.t:19-113) neither reads nor advances
    // the face field: the cells take the ordinary predictor, the faces of
    // psb the stage's initial combination
    const int rc = stage_core(c, method, dtfactor, dt, 0.0, c->slot[psa], c->slot[psb],
                              comb_slot >= 0 ? c->slot[comb_slot] : nullptr, mode, comb_a, comb_b, psb, crash_out);
    if (rc != AMRVAC_FV_OK) return rc;
    const size_t nf = (size_t)c->nb * c->nws * c->cells;
    if (mode == 1 || mode == 3) {
      const double* from = mode == 1 ? c->wsl[comb_slot] : c->wsl[psa];
      HIPCHK(hipMemcpyAsync(c->wsl[psb], from, nf * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else if (mode == 2) {  // comb_a * ws[comb_slot] + comb_b * ws[psa]
      fv::CombProg p{};
      const int ops[7] = {fv::COP_CONST, comb_slot, fv::COP_MUL, fv::COP_CONST, psa, fv::COP_MUL, fv::COP_ADD};
      p.nops = 7;
      for (int q = 0; q < 7; ++q) p.ops[q] = ops[q];
      p.k[0] = comb_a;
      p.k[1] = comb_b;
      const int gf = (int)std::min<size_t>((nf + fvc::NTHREADS - 1) / fvc::NTHREADS, 8192);
      hipLaunchKernelGGL(fv::combine_kernel, dim3(gf), dim3(fvc::NTHREADS), 0, c->stream, c->wsl[psb],
                         c->wsl[0], c->wsl[1], c->wsl[2], c->wsl[3], nf, p);
      HIPCHK(hipGetLastError());
    }
    return AMRVAC_FV_OK;
  }
  fvc::SweepArgs a = base_args(c);
  a.src = c->slot[psa];
  a.dst = c->slot[psb];
  a.base = comb_slot >= 0 ? c->slot[comb_slot] : nullptr;
  a.comb_mode = mode;
  a.comb_a = comb_a;
  a.comb_b = comb_b;
  const double qdt = dtfactor * dt;
  a.qdt = qdt;
  for (int d = 0; d < 3; ++d) a.dq[d] = -qdt / c->dx[d];
  c->crash_fresh = false;
  if (psb == 0) c->dt_valid = false;
  if (crash_out) HIPCHK(hipMemsetAsync(c->d_crash, 0, sizeof(int), c->stream));
  const bool exact = c->cfg.precision == 2;
  const double* wbase = comb_slot >= 0 ? c->wsl[comb_slot] : nullptr;
  const int bx_mode = psb == 0 ? mode : -1;
  const int mixed = c->cfg.precision == AMRVAC_FV_PREC_MIXED ? 1 : 0;
  const auto launch = c->cfg.ndim == 3 ? (exact ? fvx::launch_ct3_stage : fv::launch_ct3_stage)
                                       : (exact ? fvx::launch_ct_stage : fv::launch_ct_stage);
  hipError_t e = launch(exact ? 0 : mixed, riem, c->cfg.limiter, c->cfg.type_ct, a, c->wsl[psa], wbase, c->wsl[psb],
                        c->wsl[0], bx_mode, c->d_ctf, c->d_wsn, c->fp, c->stream);
  if (e == hipErrorInvalidValue) return fail(AMRVAC_FV_EINVAL, "no constrained-transport kernel for this scheme");
  if (e != hipSuccess) return fail(AMRVAC_FV_EHIP, std::string("ct stage: ") + hipGetErrorString(e));
  // the update kernel wrote every upper face of every block: the buffer
  // becomes psb's face field and psb's old array the next stage's buffer
  std::swap(c->wsl[psb], c->d_wsn);
  if (crash_out) HIPCHK(read_crash(c, crash_out));
  return AMRVAC_FV_OK;
}

extern "C" {

int amrvac_fv_download(amrvac_fv_ctx* c, int slot, double* w) {
  if (!c || !w || !check_slot(slot)) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (c->nb > 65535) return fail(AMRVAC_FV_EINVAL, "download supports at most 65535 blocks");
  fvc::SweepArgs a = base_args(c);
  a.src = c->slot[slot];
  const int blocks_x = (int)((c->cells_g + fvc::NTHREADS - 1) / fvc::NTHREADS);
  dim3 grid(blocks_x, c->nb);
  if (c->cfg.ndim == 2)
    hipLaunchKernelGGL((fv::fill_ghosted_kernel<2>), grid, dim3(fvc::NTHREADS), 0, c->stream, a, c->nw,
                       c->cfg.nghost, c->nxg[0], c->nxg[1], c->nxg[2], c->d_stage);
  else
    hipLaunchKernelGGL((fv::fill_ghosted_kernel<3>), grid, dim3(fvc::NTHREADS), 0, c->stream, a, c->nw,
                       c->cfg.nghost, c->nxg[0], c->nxg[1], c->nxg[2], c->d_stage);
  HIPCHK(hipGetLastError());
  const size_t n = (size_t)c->nb * c->nw * c->cells_g;
  if (c->nws == 0) {
    HIPCHK(hipMemcpyAsync(w, c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AMRVAC_FV_OK;
  }
  // constrained transport: cell variables, then the face components of
  // every block (ghost faces gathered from the neighbours like ghost cells)
  const int nwt = c->nw + c->nws;
  std::vector<double> hc(n), hf((size_t)c->nb * c->nws * c->cells_g);
  HIPCHK(hipMemcpyAsync(hc.data(), c->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  a.src = c->wsl[slot];
  if (c->cfg.ndim == 2)
    hipLaunchKernelGGL((fv::fill_ghosted_kernel<2>), grid, dim3(fvc::NTHREADS), 0, c->stream, a, c->nws,
                       c->cfg.nghost, c->nxg[0], c->nxg[1], c->nxg[2], c->d_stage);
  else
    hipLaunchKernelGGL((fv::fill_ghosted_kernel<3>), grid, dim3(fvc::NTHREADS), 0, c->stream, a, c->nws,
                       c->cfg.nghost, c->nxg[0], c->nxg[1], c->nxg[2], c->d_stage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(hf.data(), c->d_stage, hf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < c->nb; ++b)
    for (int iw = 0; iw < nwt; ++iw) {
      const double* from = iw < c->nw ? hc.data() + ((size_t)b * c->nw + iw) * c->cells_g
                                      : hf.data() + ((size_t)b * c->nws + (iw - c->nw)) * c->cells_g;
      std::memcpy(w + ((size_t)b * nwt + iw) * c->cells_g, from, c->cells_g * sizeof(double));
    }
  return AMRVAC_FV_OK;
}

int amrvac_fv_crash(amrvac_fv_ctx* c, int clear, int* crash_out) {
  if (!c || !crash_out) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->d_crash) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (!c->crash_fresh) {
    int h = 0;
    HIPCHK(read_crash(c, &h));
    c->crash_seen = h ? 1 : 0;
    c->crash_fresh = true;
  }
  *crash_out = c->crash_seen;
  if (clear) {
    HIPCHK(hipMemsetAsync(c->d_crash, 0, sizeof(int), c->stream));
    c->crash_seen = 0;
  }
  return AMRVAC_FV_OK;
}

int amrvac_fv_getdt(amrvac_fv_ctx* c, int slot, double courantpar, double* dtmin_out, double* cmax_out) {
  if (!c || !check_slot(slot)) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (slot == 0 && c->dt_valid) return getdt_reduce(c, courantpar, dtmin_out, cmax_out, c->d_dtp, c->d_red2);
  return getdt_core(c, c->slot[slot], courantpar, dtmin_out, cmax_out, c->d_red, c->d_red2);
}

int amrvac_fv_getdt_state(amrvac_fv_ctx* c, const double* u, double courantpar, double* dtmin_out,
                          double* cmax_out) {
  if (!c || !u) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (c->amr) return fail(AMRVAC_FV_EINVAL, "state arrays are for a uniform level");
  return getdt_core(c, u, courantpar, dtmin_out, cmax_out, c->d_red, c->d_red2);
}

int amrvac_fv_getdt_state_noalloc(amrvac_fv_ctx* c, const double* u, double courantpar, double* dtmin_out,
                                  double* cmax_out) {
  if (!c || !u) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (c->amr) return fail(AMRVAC_FV_EINVAL, "state arrays are for a uniform level");
  if (c->nb > kNoallocMaxBlocks)
    return fail(AMRVAC_FV_EINVAL, "amrvac_fv_getdt_state_noalloc takes at most 32768 blocks "
                                  "(its scratch array holds two partial maxima per block)");
  if (!c->scratch) return fail(AMRVAC_FV_ESTATE, "no scratch address for this device");
  // every per-block entry and the reduced triple are overwritten: nothing
  // in the scratch needs clearing between calls
  return getdt_core(c, u, courantpar, dtmin_out, cmax_out, c->scratch, c->scratch + 2 * kNoallocMaxBlocks);
}

int amrvac_fv_advect1(amrvac_fv_ctx* c, int method, double dtfactor, double dt, double cmax_global, int psa,
                      int psb, int comb_slot, double comb_a, double comb_b, int* crash_out) {
  if (!c || !check_slot(psa) || !check_slot(psb) || psa == psb) return fail(AMRVAC_FV_EINVAL, "bad slots");
  if (comb_slot >= 0 && !check_slot(comb_slot)) return fail(AMRVAC_FV_EINVAL, "bad combination slot");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  const int mode = comb_slot < 0 ? 0 : (comb_b == 0.0 ? (comb_slot == psa ? 3 : 1) : 2);
  if (c->nws > 0) return ct_stage(c, method, dtfactor, dt, psa, psb, comb_slot, mode, comb_a, comb_b, crash_out);
  return stage_core(c, method, dtfactor, dt, cmax_global, c->slot[psa], c->slot[psb],
                    comb_slot >= 0 ? c->slot[comb_slot] : nullptr, mode, comb_a, comb_b, psb, crash_out);
}

int amrvac_fv_stage(amrvac_fv_ctx* c, int method, double dtfactor, double dt, double cmax_global,
                    const double* u_in, const double* base, double comb_a, double comb_b, double* u_out) {
  if (!c || !u_in || !u_out) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (c->amr) return fail(AMRVAC_FV_EINVAL, "state arrays are for a uniform level");
  if (c->nws > 0) return fail(AMRVAC_FV_EINVAL, "constrained transport advances slots only");
  if (u_out == u_in) return fail(AMRVAC_FV_EINVAL, "u_out must not alias u_in");
  const int mode = base ? 2 : 3;
  return stage_core(c, method, dtfactor, dt, cmax_global, u_in, u_out, base, mode, comb_a, comb_b, -1, nullptr);
}

int amrvac_fv_combine(amrvac_fv_ctx* c, int dst_slot, int nops, const int* ops, int nconst, const double* consts) {
  if (!c || !check_slot(dst_slot) || nops < 1 || nops > fv::COMB_OPS || !ops || nconst < 0 ||
      nconst > fv::COMB_CONSTS || (nconst > 0 && !consts))
    return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  fv::CombProg p{};
  p.nops = nops;
  int depth = 0, nk = 0;
  for (int q = 0; q < nops; ++q) {
    const int op = ops[q];
    if ((op >= 0 && op <= 3) || op == fv::COP_CONST) {
      if (op == fv::COP_CONST && nk >= nconst) return fail(AMRVAC_FV_EINVAL, "combination: too few constants");
      nk += op == fv::COP_CONST;
      if (++depth > fv::COMB_DEPTH) return fail(AMRVAC_FV_EINVAL, "combination: stack deeper than 8");
    } else if (op == fv::COP_NEG) {
      if (depth < 1) return fail(AMRVAC_FV_EINVAL, "combination: stack underflow");
    } else if (op >= fv::COP_ADD && op <= fv::COP_DIV) {
      if (depth < 2) return fail(AMRVAC_FV_EINVAL, "combination: stack underflow");
      --depth;
    } else {
      return fail(AMRVAC_FV_EINVAL, "combination: unknown operation");
    }
    p.ops[q] = op;
  }
  if (depth != 1 || nk != nconst) return fail(AMRVAC_FV_EINVAL, "combination: malformed program");
  if (dst_slot == 0) c->dt_valid = false;
  for (int i = 0; i < nconst; ++i) p.k[i] = consts[i];
  const size_t n = (size_t)c->nb * c->nw * c->cells;
  const int grid = (int)std::min<size_t>((n + fvc::NTHREADS - 1) / fvc::NTHREADS, 8192);
  hipLaunchKernelGGL(fv::combine_kernel, dim3(grid), dim3(fvc::NTHREADS), 0, c->stream, c->slot[dst_slot],
                     c->slot[0], c->slot[1], c->slot[2], c->slot[3], n, p);
  HIPCHK(hipGetLastError());
  if (c->nws > 0) {  // the face field takes the same combination (src/mod_advance.t:119-291)
    const size_t nf = (size_t)c->nb * c->nws * c->cells;
    const int gf = (int)std::min<size_t>((nf + fvc::NTHREADS - 1) / fvc::NTHREADS, 8192);
    hipLaunchKernelGGL(fv::combine_kernel, dim3(gf), dim3(fvc::NTHREADS), 0, c->stream, c->wsl[dst_slot],
                       c->wsl[0], c->wsl[1], c->wsl[2], c->wsl[3], nf, p);
    HIPCHK(hipGetLastError());
  }
  return AMRVAC_FV_OK;
}

int amrvac_fv_copy(amrvac_fv_ctx* c, int dst, int src) {
  if (!c || !check_slot(dst) || !check_slot(src)) return fail(AMRVAC_FV_EINVAL, "bad slots");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (dst == src) return AMRVAC_FV_OK;
  if (dst == 0) c->dt_valid = false;
  HIPCHK(hipMemcpyAsync(c->slot[dst], c->slot[src], (size_t)c->nb * c->nw * c->cells * sizeof(double),
                        hipMemcpyDeviceToDevice, c->stream));
  if (c->nws > 0)
    HIPCHK(hipMemcpyAsync(c->wsl[dst], c->wsl[src], (size_t)c->nb * c->nws * c->cells * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
  return AMRVAC_FV_OK;
}

int amrvac_fv_volume_average(amrvac_fv_ctx* c, int slot, double* modes) {
  if (!c || !modes || !check_slot(slot)) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  hipLaunchKernelGGL(fv::volsum_kernel, dim3(c->nb), dim3(fvc::NTHREADS), 0, c->stream, c->slot[slot], c->nb,
                     c->nw, c->cells, c->d_red);
  HIPCHK(hipGetLastError());
  std::vector<double> h((size_t)c->nb * 2 * c->nw);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_red, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // src/io/mod_input_output.t:2694-2729: volume-weighted sums; the cells of
  // a block share one volume (dx of its level)
  std::vector<double> dvol(c->nb);
  double vol = 0.0;
  for (int b = 0; b < c->nb; ++b) {
    double v = 1.0;
    const double sc = c->amr ? std::ldexp(1.0, -(c->h_level[b] - 1)) : 1.0;
    for (int d = 0; d < c->cfg.ndim; ++d) v = v * (c->dx[d] * sc);
    dvol[b] = v;
    vol += v * (double)c->cells;
  }
  for (int p = 0; p < 2; ++p)
    for (int iw = 0; iw < c->nw; ++iw) {
      double s = 0.0;
      for (int b = 0; b < c->nb; ++b) s += dvol[b] * h[((size_t)b * 2 + p) * c->nw + iw];
      modes[p * c->nw + iw] = s / vol;
    }
  return AMRVAC_FV_OK;
}

int amrvac_fv_slot_pointer(amrvac_fv_ctx* c, int slot, void** dptr, long long* nbytes) {
  if (!c || !check_slot(slot) || !dptr) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  *dptr = c->slot[slot];
  if (nbytes) *nbytes = (long long)((size_t)c->nb * c->nw * c->cells * sizeof(double));
  // writes through the pointer bypass the context: ps's CFL partials are
  // no longer kept by the stages of this context
  if (slot == 0) {
    c->slot0_exposed = true;
    c->dt_valid = false;
  }
  return AMRVAC_FV_OK;
}


int amrvac_fv_set_mesh_amr(amrvac_fv_ctx* c, int nblocks, const double* dx1, const int* level, const int* ntype,
                           const int* nbr, const int* nchild, const int* icpos, const double* xlohi, const int* bc) {
  if (!c || !dx1 || !level || !ntype || !nbr || !nchild || !icpos || !xlohi || !bc)
    return fail(AMRVAC_FV_EINVAL, "null argument");
  if (nblocks < 1) return fail(AMRVAC_FV_EINVAL, "nblocks must be positive");
  if (c->nws > 0) return fail(AMRVAC_FV_EINVAL, "constrained transport is provided on a uniform level");
  if (c->bx % 2 != 0 || (c->cfg.ndim == 3 && c->bnz % 2 != 0))
    return fail(AMRVAC_FV_EINVAL, "an AMR mesh needs even block_nx");
  const int nd = c->cfg.ndim, nnb = c->nnb, nch = 1 << nd;
  for (int b = 0; b < nblocks; ++b) {
    if (level[b] < 1 || level[b] > fvc::MAXLEV) return fail(AMRVAC_FV_EINVAL, "level out of range (1..8)");
    for (int d = 0; d < nd; ++d)
      if (icpos[b * nd + d] != 1 && icpos[b * nd + d] != 2) return fail(AMRVAC_FV_EINVAL, "icpos must be 1 or 2");
    for (int k = 0; k < nnb; ++k) {
      const int t = ntype[b * nnb + k], m = nbr[b * nnb + k];
      if (k == (nnb - 1) / 2) {
        if (m != b) return fail(AMRVAC_FV_EINVAL, "neighbour table centre entry must be the block itself");
        continue;
      }
      if (t < 1 || t > 4) return fail(AMRVAC_FV_EINVAL, "neighbour type must be 1..4");
      if ((t == 2 || t == 3) && (m < 0 || m >= nblocks)) return fail(AMRVAC_FV_EINVAL, "neighbour index out of range");
      if (t == 3 && level[m] != level[b]) return fail(AMRVAC_FV_EINVAL, "sibling on another level");
      if (t == 2 && level[m] != level[b] - 1) return fail(AMRVAC_FV_EINVAL, "coarse neighbour not one level up");
      if (t == 4) {
        int i[3] = {0, 0, 0}, kk = k;
        for (int d = 0; d < nd; ++d) { i[d] = kk % 3 - 1; kk /= 3; }
        for (int j = 0; j < nch; ++j) {
          bool adjacent = true;  // children of the neighbour node touching block b
          for (int d = 0; d < nd; ++d) {
            const int ic = 1 + ((j >> d) & 1);
            if ((i[d] == -1 && ic != 2) || (i[d] == 1 && ic != 1)) adjacent = false;
          }
          if (!adjacent) continue;
          const int f = nchild[((size_t)b * nnb + k) * nch + j];
          if (f < 0 || f >= nblocks || level[f] != level[b] + 1)
            return fail(AMRVAC_FV_EINVAL, "fine neighbour child index or level out of range");
        }
      }
    }
  }
  for (int s = 0; s < 2 * nd; ++s)
    if (bc[s] != 2 && bc[s] != 3 && bc[s] != 4 && bc[s] != 5)
      return fail(AMRVAC_FV_EINVAL, "boundary type must be cont, symm, asymm or periodic");
  for (int d = 0; d < nd; ++d)
    if (!(dx1[d] > 0.0)) return fail(AMRVAC_FV_EINVAL, "dx must be positive");
  HIPCHK(hipStreamSynchronize(c->stream));
  // a new AMR level over an AMR level keeps the buffers that still fit
  if (c->amr) reset_level(c);
  else free_level(c);
  // virtual blocks and the sweep's neighbour table
  std::vector<int> snbr((size_t)nblocks * nnb);
  std::vector<int4> rt, pt;
  std::vector<int2> pt2;
  int nv = 0;
  for (int b = 0; b < nblocks; ++b)
    for (int k = 0; k < nnb; ++k) {
      const int t = ntype[b * nnb + k];
      if (k == (nnb - 1) / 2) { snbr[(size_t)b * nnb + k] = b; continue; }
      if (t == 3) { snbr[(size_t)b * nnb + k] = nbr[b * nnb + k]; continue; }
      if (t == 1) { snbr[(size_t)b * nnb + k] = -1; continue; }
      const int v = nblocks + nv++;
      snbr[(size_t)b * nnb + k] = v;
      if (t == 4) {
        rt.push_back(make_int4(b, k, v, 0));
      } else {
        int i[3] = {0, 0, 0}, kk = k;
        for (int d = 0; d < nd; ++d) { i[d] = kk % 3 - 1; kk /= 3; }
        int pic = 0, pdir = 0;
        for (int d = 0; d < nd; ++d) {
          const int ic = icpos[b * nd + d];
          // direction from the coarse neighbour to the block's parent
          const int q = ic - 1 + i[d];
          const int iC = -(q >= 0 ? q / 2 : -((-q + 1) / 2));
          pic |= ic << (2 * d);
          pdir |= (iC + 1) << (2 * d);
        }
        pt.push_back(make_int4(b, k, v, nbr[b * nnb + k]));
        pt2.push_back(make_int2(pic, pdir));
      }
    }
  int rc = alloc_level(c, nblocks, nblocks + nv, dx1, snbr.data(), bc);
  if (rc != AMRVAC_FV_OK) return rc;
  c->amr = true;
  c->nv = nv;
  c->nrt = (int)rt.size();
  c->npt = (int)pt.size();
  // refluxing registers on faces with a level jump, coarse blocks to fix
  c->nface = c->bx * (c->bnz > c->bx ? c->bnz : c->bx);
  std::vector<int> fregoff((size_t)nblocks * 2 * nd, -1), ftask;
  int nreg = 0;
  for (int b = 0; b < nblocks; ++b) {
    bool fixes = false;
    for (int d = 0; d < nd; ++d)
      for (int sd = 0; sd < 2; ++sd) {
        int v[3] = {0, 0, 0};
        v[d] = sd == 0 ? -1 : 1;
        int k = 0, m = 1;
        for (int e = 0; e < nd; ++e) { k += (v[e] + 1) * m; m *= 3; }
        const int t = ntype[b * nnb + k];
        if (t == 2 || t == 4) fregoff[(size_t)b * 2 * nd + 2 * d + sd] = (nreg++) * c->nw * c->nface;
        if (t == 4) fixes = true;
      }
    if (fixes) ftask.push_back(b);
  }
  c->nft = (int)ftask.size();
  c->h_level.assign(level, level + nblocks);
  std::vector<int> lev0(nblocks);
  for (int b = 0; b < nblocks; ++b) lev0[b] = level[b] - 1;
  std::vector<int> vtype(ntype, ntype + (size_t)nblocks * nnb), vchild(nchild, nchild + (size_t)nblocks * nnb * nch);
  std::vector<double> xlo((size_t)nblocks * nd);
  for (int b = 0; b < nblocks; ++b)
    for (int d = 0; d < nd; ++d) xlo[(size_t)b * nd + d] = xlohi[((size_t)b * nd + d) * 2];
  bool ok = upload_vec(c, &c->d_level, &c->level_cap, lev0) && upload_vec(c, &c->d_ntype, &c->ntype_cap, vtype) &&
            upload_vec(c, &c->d_nchild, &c->nchild_cap, vchild) && upload_vec(c, &c->d_rtask, &c->rtask_cap, rt) &&
            upload_vec(c, &c->d_ptask, &c->ptask_cap, pt) && upload_vec(c, &c->d_ptask2, &c->ptask2_cap, pt2) &&
            upload_vec(c, &c->d_ftask, &c->ftask_cap, ftask) &&
            upload_vec(c, &c->d_fregoff, &c->fregoff_cap, fregoff) && upload_vec(c, &c->d_xlo, &c->xlo_cap, xlo);
  if (ok && nreg > 0) ok = ensure_buf(&c->d_freg, &c->freg_cap, (size_t)nreg * c->nw * c->nface);
  c->nreg = nreg;
  if (!ok) {
    free_level(c);
    return fail(AMRVAC_FV_ENOMEM, "AMR table allocation failed");
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRVAC_FV_OK;
}

int amrvac_fv_errest(amrvac_fv_ctx* c, int slot, int nwt, const int* iws, const double* weights, double wavefilter,
                     double threshold, double coarsen_threshold, unsigned char* flags) {
  if (!c || !check_slot(slot) || nwt < 1 || nwt > 64 || !iws || !weights || !flags)
    return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  for (int t = 0; t < nwt; ++t)
    if (iws[t] < 0 || iws[t] >= c->nw) return fail(AMRVAC_FV_EINVAL, "errest variable out of range");
  if (c->cfg.nghost < 2) return fail(AMRVAC_FV_EINVAL, "the Löhner estimate reads two ghost layers");
  if (!c->d_iws) {
    if (hipMalloc(&c->d_iws, 64 * sizeof(int)) != hipSuccess || hipMalloc(&c->d_wts, 64 * sizeof(double)) != hipSuccess)
      return fail(AMRVAC_FV_ENOMEM, "errest buffers");
  }
  // flags, then (4-byte aligned) the copy of the small-value flag
  const size_t nbf = ((size_t)c->nb + 3) & ~(size_t)3;
  if (!ensure_buf(&c->d_flags, &c->flags_cap, nbf + sizeof(int)) || !ensure_pin(c, kPinFlags + nbf + sizeof(int)))
    return fail(AMRVAC_FV_ENOMEM, "errest flags");
  // the variables and weights go up when they change (a run passes the
  // same ones every time)
  const std::vector<int> iv(iws, iws + nwt);
  const std::vector<double> wv(weights, weights + nwt);
  if (iv != c->h_iws || wv != c->h_wts) {
    HIPCHK(hipMemcpyAsync(c->d_iws, iws, nwt * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_wts, weights, nwt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // pageable sources: the copies have read them when the calls return
    c->h_iws = iv;
    c->h_wts = wv;
  }
  fvc::SweepArgs a = base_args(c);
  a.src = c->slot[slot];
  const bool tiled2 = c->cfg.ndim == 2 && c->bx * c->bx == fvc::NTHREADS && c->bx + 4 == fv::kLohnerT;
  const bool tiled3 = c->cfg.ndim == 3 && (c->bx + 4) * (c->bx + 4) * (c->bnz + 4) <= fv::kLohner3Tile &&
                      c->bx * c->bx * c->bnz <= fv::kLohner3Cells * fvc::NTHREADS;
  const bool tiled = tiled2 || tiled3;
  int* crash_copy = reinterpret_cast<int*>(c->d_flags + nbf);
  if (tiled2)
    hipLaunchKernelGGL(fv::lohner2_kernel, dim3(c->nb), dim3(fvc::NTHREADS), 0, c->stream, a, c->nw, nwt, c->d_iws,
                       c->d_wts, wavefilter, threshold, coarsen_threshold, c->d_flags, (const int*)c->d_crash,
                       crash_copy);
  else if (tiled3)
    hipLaunchKernelGGL(fv::lohner3_kernel, dim3(c->nb), dim3(fvc::NTHREADS), 0, c->stream, a, c->nw, nwt, c->d_iws,
                       c->d_wts, wavefilter, threshold, coarsen_threshold, c->d_flags, (const int*)c->d_crash,
                       crash_copy);
  else if (c->cfg.ndim == 2)
    return fail(AMRVAC_FV_EINVAL, "the 2-D Löhner estimate is built for 16^2 blocks");
  else
    hipLaunchKernelGGL((fv::lohner_kernel<3>), dim3(c->nb), dim3(fvc::NTHREADS), 0, c->stream, a, c->nw, nwt,
                       c->d_iws, c->d_wts, wavefilter, threshold, coarsen_threshold, c->d_flags);
  HIPCHK(hipGetLastError());
  // the small-value flag rides the same synchronisation (amrvac_fv_crash
  // then reads it without another one): copied behind the flags by the
  // tiled kernel, else on its own
  if (tiled) {
    HIPCHK(hipMemcpyAsync(c->h_pin + kPinFlags, c->d_flags, nbf + sizeof(int), hipMemcpyDeviceToHost, c->stream));
  } else {
    HIPCHK(hipMemcpyAsync(c->h_pin + kPinFlags, c->d_flags, (size_t)c->nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_pin + kPinFlags + nbf, c->d_crash, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  memcpy(flags, c->h_pin + kPinFlags, (size_t)c->nb);
  const int h = *reinterpret_cast<const int*>(c->h_pin + kPinFlags + nbf);
  c->crash_seen = h ? 1 : 0;
  c->crash_fresh = true;
  return AMRVAC_FV_OK;
}

int amrvac_fv_regrid(amrvac_fv_ctx* c, int nblocks, const int* plan, const double* dx1, const int* level,
                     const int* ntype, const int* nbr, const int* nchild, const int* icpos, const double* xlohi,
                     const int* bc) {
  if (!c || !plan || nblocks < 1) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0] || !c->amr) return fail(AMRVAC_FV_ESTATE, "regrid needs an AMR mesh");
  const int nd = c->cfg.ndim, S = 2 + (1 << nd), nold = c->nb;
  for (int b = 0; b < nblocks; ++b) {
    const int* p = plan + (size_t)b * S;
    if (p[0] == 0 || p[0] == 2) {
      if (p[1] < 0 || p[1] >= nold) return fail(AMRVAC_FV_EINVAL, "regrid plan: old block out of range");
      if (p[0] == 2 && (p[S - 1] < 0 || p[S - 1] >= (1 << nd)))
        return fail(AMRVAC_FV_EINVAL, "regrid plan: child position out of range");
    } else if (p[0] == 1) {
      for (int j = 0; j < (1 << nd); ++j)
        if (p[1 + j] < 0 || p[1 + j] >= nold) return fail(AMRVAC_FV_EINVAL, "regrid plan: old child out of range");
    } else {
      return fail(AMRVAC_FV_EINVAL, "regrid plan: kind must be 0 (copy), 1 (coarsen) or 2 (refine)");
    }
  }
  const size_t cells = (size_t)c->nw * c->cells;
  if (!ensure_buf(&c->d_remap, &c->remap_cap, (size_t)nblocks * cells) ||
      !ensure_buf(&c->d_plan, &c->plan_cap, (size_t)nblocks * S))
    return fail(AMRVAC_FV_ENOMEM, "regrid buffers");
  HIPCHK(hipMemcpyAsync(c->d_plan, plan, (size_t)nblocks * S * sizeof(int), hipMemcpyHostToDevice, c->stream));
  fvc::SweepArgs a = base_args(c);
  a.src = c->slot[0];
  if (nd == 2)
    hipLaunchKernelGGL((fv::remap_kernel<2>), dim3(nblocks), dim3(fvc::NTHREADS), 0, c->stream, a, c->nw, c->d_plan,
                       c->d_remap);
  else
    hipLaunchKernelGGL((fv::remap_kernel<3>), dim3(nblocks), dim3(fvc::NTHREADS), 0, c->stream, a, c->nw, c->d_plan,
                       c->d_remap);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  // the new connectivity (frees the old level), then the new leaves' data and
  // their ghost cells
  const int rc = amrvac_fv_set_mesh_amr(c, nblocks, dx1, level, ntype, nbr, nchild, icpos, xlohi, bc);
  if (rc != AMRVAC_FV_OK) return rc;
  HIPCHK(hipMemcpyAsync(c->slot[0], c->d_remap, (size_t)nblocks * cells * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  return getbc_amr(c, 0);
}

int amrvac_fv_getbc(amrvac_fv_ctx* c, int slot) {
  if (!c || !check_slot(slot)) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  return getbc_amr(c, slot);
}

int amrvac_fv_set_fix_conserve(amrvac_fv_ctx* c, int on) {
  if (!c) return fail(AMRVAC_FV_EINVAL, "null context");
  c->fix_conserve = on != 0;
  return AMRVAC_FV_OK;
}


int amrvac_fv_adjoint1(amrvac_fv_ctx* c, int method, double dtfactor, double dt, double cmax_global,
                       const double* u_in, const double* lam_out, double coef, const double* add_a, double ca,
                       const double* add_b, double cb, double* lam_in) {
  if (!c || !u_in || !lam_out || !lam_in) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  if (c->nws > 0) return fail(AMRVAC_FV_EINVAL, "adjoint: no constrained transport");
  if (c->bx != 16) return fail(AMRVAC_FV_EINVAL, "the adjoint stage needs block_nx = 16");
  if (lam_in == lam_out || lam_in == u_in) return fail(AMRVAC_FV_EINVAL, "lam_in must not alias lam_out or u_in");
  if (c->cfg.physics == 2 && ((c->cfg.divb != 0 && c->cfg.divb != 1) || c->cfg.ndir != 3))
    return fail(AMRVAC_FV_EINVAL, "the MHD adjoint stage is provided for ndir 3, without divergence control "
                                  "or with GLM");
  const bool glm = c->cfg.physics == 2 && c->cfg.divb == 1;
  if (glm && !(cmax_global > 0.0)) return fail(AMRVAC_FV_EINVAL, "GLM needs cmax_global > 0");
  if (c->cfg.physics != 1 && c->cfg.physics != 2) return fail(AMRVAC_FV_EINVAL, "unknown physics");
  if (c->amr) return fail(AMRVAC_FV_EINVAL, "the adjoint stage needs a uniform level");
  for (int sd = 0; sd < 2 * c->cfg.ndim; ++sd)
    if (c->bc[sd] != 5) return fail(AMRVAC_FV_EINVAL, "the adjoint stage needs periodic boundaries");
  if (c->cfg.boundspeed != 1) return fail(AMRVAC_FV_EINVAL, "the adjoint stage needs Einfeldt bounds");
  if (c->cfg.limiter != 1 && c->cfg.limiter != 5 && c->cfg.limiter != 9)
    return fail(AMRVAC_FV_EINVAL, "the adjoint needs minmod, vanleer or cada3");
  int riem;
  if (method_to_riem(method, riem) != 0 || riem == fv::RIEM_HANCOCK || riem == fv::RIEM_HLLD)
    return fail(AMRVAC_FV_EINVAL, "adjoint: flux scheme must be hll, hllc or tvdlf");
  if (riem == fv::RIEM_HLLC && c->cfg.physics != 1)
    return fail(AMRVAC_FV_EINVAL, "adjoint: hllc is provided for hd");
  if (c->cfg.ndim == 3 && !c->d_adjG) {
    const size_t n = (size_t)c->nw * c->nb * c->cells;
    if (hipMalloc(&c->d_adjG, n * sizeof(double)) != hipSuccess) return fail(AMRVAC_FV_ENOMEM, "adjoint buffer");
  }
  fvc::SweepArgs a = base_args(c);
  a.src = u_in;
  const double qdt = dtfactor * dt;
  a.qdt = qdt;
  for (int d = 0; d < 3; ++d) a.dq[d] = -qdt / c->dx[d];
  a.ch = cmax_global;
  fvc::AdjArgs g{lam_out, coef, add_a, ca, add_b, cb, lam_in, c->d_adjG};
  const int mixed = c->cfg.precision == AMRVAC_FV_PREC_MIXED ? 1 : 0;
  hipError_t e = fv::launch_adjoint_stage(c->cfg.ndim, c->cfg.physics, glm ? 1 : 0, mixed, riem, c->cfg.limiter, a,
                                          c->fp, g, c->stream);
  if (e == hipErrorInvalidValue) return fail(AMRVAC_FV_EINVAL, "no adjoint kernel for this configuration");
  if (e != hipSuccess) return fail(AMRVAC_FV_EHIP, std::string("adjoint kernels: ") + hipGetErrorString(e));
  return AMRVAC_FV_OK;
}


int amrvac_fv_slot_copy(amrvac_fv_ctx* c, int slot, void* dptr, int to_slot) {
  if (!c || !check_slot(slot) || !dptr) return fail(AMRVAC_FV_EINVAL, "bad argument");
  if (!c->slot[0]) return fail(AMRVAC_FV_ESTATE, "no mesh");
  // a slot of a constrained-transport context also holds its face field,
  // which a cell-only copy would leave stale beside the restored cells
  if (c->nws > 0) return fail(AMRVAC_FV_EINVAL, "slot copies move cell variables only: not with constrained transport");
  if (to_slot && slot == 0) c->dt_valid = false;
  const size_t n = (size_t)c->nb * c->nw * c->cells * sizeof(double);
  if (to_slot) HIPCHK(hipMemcpyAsync(c->slot[slot], dptr, n, hipMemcpyDeviceToDevice, c->stream));
  else HIPCHK(hipMemcpyAsync(dptr, c->slot[slot], n, hipMemcpyDeviceToDevice, c->stream));
  return AMRVAC_FV_OK;
}

int amrvac_fv_sync(amrvac_fv_ctx* c) {
  if (!c) return fail(AMRVAC_FV_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRVAC_FV_OK;
}

int amrvac_fv_sweep_timing(amrvac_fv_ctx* c, int reset, double* avg_ms, int* count) {
  if (!c) return fail(AMRVAC_FV_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  for (size_t i = 0; i < c->nev; ++i) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->events[i].first, c->events[i].second));
    tot += ms;
  }
  if (avg_ms) *avg_ms = c->nev ? tot / (double)c->nev : 0.0;
  if (count) *count = (int)c->nev;
  if (reset) {
    c->nev = 0;
    c->timing = true;
  }
  return AMRVAC_FV_OK;
}

}  // extern "C"
