// Internal declarations shared by the operator, Krylov and API translation units.
#pragma once
#include "cgm_common.h"

namespace cgm {

enum {
  kModeForward = 0,      // y = x - G(chi x)
  kModeAdjoint = 1,      // y = x - chi G(x)
  kModeGreen = 2,        // y = G(x)
  kModeResidual = 3,     // y = b - (x - G(chi x))
  kModeResidualAdj = 4   // y = b - (x - chi G(x))
};

// Per-system strides of a batched launch: system k = blockIdx.y reads and
// writes at base + k * stride.  All zero (and gridDim.y = 1) for one system.
struct BatchStrides {
  long x = 0;    // input vector x (elements)
  long chi = 0;  // contrast (elements)
  long b = 0;    // right-hand side b (elements)
  long y = 0;    // output vector y (elements)
  long T = 0;    // FFT scratch (elements)
  long st = 0;   // DevState (bytes): norm and done pointers
};

template <typename P>
__host__ __device__ inline P* boff(P* p, long bytes) {
  return p ? reinterpret_cast<P*>(reinterpret_cast<uintptr_t>(p) + bytes) : p;
}

// The N x M scratch of a matvec is stored in column panels of kPW columns:
// element (r, c) sits at ((c / kPW) * N + r) * kPW + c % kPW.  Row passes
// (1 and 3) then move runs of kPW contiguous values per row, and a column
// pass work-group with kPW columns reads and writes one contiguous block.
constexpr int kPW = 8;
__host__ __device__ inline long tpos(int r, int c, int N) {
  return ((long)(c / kPW) * N + r) * kPW + (c % kPW);
}

struct Plan {
  int N, M;
  int rows1, rows3;  // rows per work-group in passes 1 and 3
  int cols2;         // columns per work-group in pass 2
  int wide1_batch, wide2_batch, wide3_batch;  // M = 1024, complex64: launches of at
                     // least this many vectors take the register-resident
                     // pass 1 / 2 / 3 (0: never)
};

int plan_init(Plan& P, int N);
int fft_threads(int M);
int choose_embedding(int N);
int pass_max_threads(int M, bool fp64);

template <typename T>
int launch_pass1(const Plan& P, hipStream_t st, const cplx<T>* x, cplx<T>* xw, const T* chi,
                 const double* norm_ptr, int mode, cplx<T>* Tm, const cplx<T>* tw, const int* done,
                 int B = 1, const BatchStrides& bs = BatchStrides());
template <typename T, typename G>
int launch_pass2(const Plan& P, hipStream_t st, cplx<T>* Tm, const G* gq, const cplx<T>* tw,
                 const int* done, int B = 1, const BatchStrides& bs = BatchStrides());
template <typename T>
int launch_pass3(const Plan& P, hipStream_t st, const cplx<T>* Tm, const cplx<T>* x, const T* chi,
                 const cplx<T>* b, cplx<T>* y, int mode, const cplx<T>* tw, const int* done,
                 int B = 1, const BatchStrides& bs = BatchStrides());

int spectrum_from_block(const Plan& P, hipStream_t st, const double2* blk, const double2* tw64,
                        float2* gq32, double2* gq64);
int green_block(hipStream_t st, double2* blk, int N, int kind, double k, double sr, double sc);
int gprep_block(hipStream_t st, const float2* gprep, double2* blk, int N);

// Solver state in device memory (one per context).
struct DevState {
  double beta;     // ||r|| at the start of the current cycle
  double bnorm;    // ||b||
  double hn;       // last ||w|| (H[i+1, i])
  double tol_abs;  // tol * ||b||
  double res_est;  // latest |g_{i+1}|
  double norm_out; // result slot of the stand-alone norm kernel
  double unorm;    // ||u_{j+1}|| from the flexible fused step (the network's input scale)
  int done;        // 1: remaining launches of this cycle are no-ops
  int k;           // valid Arnoldi columns of this cycle
  int status;      // 0 or CGM_ERROR_NONFINITE
  int k_last;      // k of the last back substitution (kept across the next cycle's reset)
  unsigned tickets[8];
};

enum { kTicketDots1 = 0, kTicketDots2 = 1, kTicketNorm = 2, kTicketInit = 3, kTicketMisc = 4 };

constexpr int kMaxRestart = 512;
constexpr int kFusedMaxRestart = 64;  // longest cycle of the fused lagged-CGS2 kernels
constexpr int kKB = 1024;  // threads per block of the streaming Krylov kernels

struct KrylovWs {
  int m = 0;              // restart this workspace was sized for
  int B = 1;              // systems (lockstep batch)
  long sBasis = 0;        // per-system strides (elements): V
  long sSmall = 0;        //   every small FP64 array (H, R, g, cs, sn, y, h1, h2, h1p, a1, a2, scal)
  long sPart = 0;         //   part
  long sPartn = 0;        //   partn
  long sVec = 0;          //   Z (and half of U)
  int G = 0;              // blocks of the reduction kernels
  long n = 0;
  float2* V = nullptr;    // (m+1) * n basis
  double2* h1 = nullptr;  // first CGS pass coefficients (m+1)
  double2* h2 = nullptr;  // second CGS pass coefficients (m+1)
  double* cs = nullptr;   // Givens cosines (m)
  double2* sn = nullptr;  // Givens sines (m)
  double2* g = nullptr;   // rotated right-hand side (m+1)
  double2* R = nullptr;   // upper triangular factor, row-major m x m
  double2* y = nullptr;   // least-squares solution (m)
  double2* part = nullptr;  // reduction partials (2m+4) * max(G, fused blocks)
  double* partn = nullptr;  // norm partials G
  // lagged-CGS2 (fused) path
  double2* H = nullptr;     // (m+1) x m Hessenberg, row-major
  double2* h1p = nullptr;   // (m+1)
  double2* a1 = nullptr;    // (m+1)
  double2* a2 = nullptr;    // (m+1)
  double2* scal = nullptr;  // 4
  float2* U = nullptr;      // 2 * n (u ping-pong)
  float2* Z = nullptr;      // n
};

struct FusedLaunch {
  int N, m, ld, j, mode;  // m: steps of this cycle; ld: leading dimension of H and R
  int flex = 0;           // flexible (right-preconditioned) steps: the matvec input is Xin = z_j
  int small = 0;          // 256-thread work-groups where available (see fused_shape)
  const float2* Xin = nullptr;
  long sXin = 0;            // per-system stride of Xin (elements)
  double* partn = nullptr;  // norm partials (flex)
  int B = 1;              // systems; per-system strides below (elements, DevState in bytes)
  long sChi = 0, sBasis = 0, sU = 0, sVec = 0, sT = 0, sPart = 0, sSmall = 0, sSt = 0, sHist = 0;
  long sPartn = 0;
  const float* chi;
  float2* V;
  const float2* Uj;
  float2* Unext;
  float2* Z;
  float2* T;
  const float2* tw;
  double2* part;
  double2* H;
  double2* Rm;
  double* cs;
  double2* sn;
  double2* g;
  double2* h1p;
  double2* a1;
  double2* a2;
  double2* scal;
  DevState* st;
  double* hist;
};
int gmres_fixed_cycle(cgm_context* c, const float2* gq, const float* chi, const float2* b, float2* x, int m,
                      int adjoint, double tolerance, double* hist, cgm_gmres_result* res);
bool fused_supported(int N, int restart);
int fused_blocks(int N);
int launch_fused(hipStream_t st, int which, const FusedLaunch& L);

// Krylov launches (cgm_krylov.hip)
int krylov_alloc(KrylovWs& ws, long n, int m, hipStream_t st, int B = 1);
void krylov_free(KrylovWs& ws);
// Batched launches take per-system strides of r (elements) and of the
// DevState array (bytes) and the number of systems of the call (at most
// ws.B); one system by default.
int launch_norm_init(hipStream_t st, const KrylovWs& ws, const float2* r, DevState* s, int what,
                     double tol, long sr = 0, long sst = 0, int nsys = 1);
int launch_cgs2_step(hipStream_t st, const KrylovWs& ws, int i, DevState* s, double* hist,
                     int part = -1);
// x += basis y (basis = ws.V when null; FGMRES passes its Z basis)
int launch_backsub_update(hipStream_t st, const KrylovWs& ws, DevState* s, float2* x,
                          const float2* basis = nullptr, long sx = 0, long sst = 0, int nsys = 1);
int launch_norm64(hipStream_t st, const double2* v, long n, double* partn, int G, DevState* s);

// Long-recurrence flexible GMRES (cgm_longrec.hip): restart up to
// kMaxRestartLR, the whole basis in HBM, lagged CGS2 with two basis passes
// per step.
constexpr int kMaxRestartLR = 16384;
struct LrWs {
  int m = 0, B = 0;
  long n = 0;
  int nchunks = 0, nuchunks = 0;
  long sPart = 0, sSmall = 0;
  float2* Q = nullptr;     // [B][m+1][n]
  float2* Z = nullptr;     // [B][m][n]
  float2* U = nullptr;     // [B][2][n]
  float2* W = nullptr;     // [B][n]
  double2* part = nullptr;
  double2 *s = nullptr, *t = nullptr, *h1p = nullptr, *g = nullptr, *y = nullptr, *sn = nullptr;
  double* cs = nullptr;
  double2 *a1 = nullptr, *a2 = nullptr;
  double* scal = nullptr;
  double2* R = nullptr;    // [B][m][m] column-major
  double* partn = nullptr;
  // FP64 products w = A z_j of the current step (the context's buffer, n per
  // system) when the products are FP64; null: w is the complex64 W.
  const double2* W64 = nullptr;
};
// Live timing of the streaming step kernels over a solve (0 projection,
// 1 update, 2 tile-inverse application): HIP events around each launch on
// the solver's stream, read back every kRing steps; summed durations,
// algorithmic bytes and launch counts.  `cur` is the ring slot of the step
// whose preconditioner is being enqueued (-1 outside a long cycle); the tile
// application records into it.
struct LrTiming {
  static constexpr int kRing = 32;
  static constexpr int kKernels = 3;
  bool on = false, created = false;
  int cur = -1;
  hipEvent_t ev[2 * kKernels * kRing];
  double slot_bytes[kKernels][kRing];
  bool slot_has[kKernels][kRing];
  double ms[kKernels] = {0.0, 0.0, 0.0}, bytes[kKernels] = {0.0, 0.0, 0.0};
  long launches[kKernels] = {0, 0, 0};
};
int lr_alloc(LrWs& w, long n, int m, int B);
void lr_free(LrWs& w);
// Device bytes of a long-recurrence workspace.
size_t lr_bytes(long n, int m, int B);
// step j: (j, u_j, its per-system stride, z_j or w output, its stride)
using LrStepFn = int (*)(void* ctx, int j, const float2* in, long sin, float2* out, long sout);
struct LrStep {
  LrStepFn fn;
  void* ctx;
  int operator()(int j, const float2* in, long sin, float2* out, long sout) const {
    return fn(ctx, j, in, sin, out, sout);
  }
};
int lr_cycle(LrWs& w, hipStream_t st, DevState* dst, DevState* hst, long sst, int B, int m, double* hist,
             long shist, float2* x, long sx, const LrStep& precond, const LrStep& matvec, LrTiming* tm = nullptr);
int lr_scale(hipStream_t st, const float2* u, long su, float2* z, long sz, long n, const DevState* dst, long sst,
             int B);
// Device address of the allocation-free matvec's scratch array on `device`
// (cgm_noalloc.hip), taken once per device; the device must be current.
int noalloc_init(int device);
}  // namespace cgm

struct cgm_context;
struct cgm_fno;
struct cgm_tiles;
namespace cgm {
// FNO preconditioner (cgm_fno.hip): z = P(v) for B vectors; with nrm the
// input is v / *nrm, written back to vw when given (v is left unchanged
// otherwise).  Every launch is skipped once *done.
int fno_forward(cgm_fno* f, hipStream_t st, const float* chi, const float2* v, const double* nrm, float2* vw,
                float2* z, int B, bool train, const int* done);
// The same with per-sample strides (elements; the norm's in bytes): sample k
// reads chi + k sChi (sChi = 0: one shared map), v + k sV, *(nrm + k sNrm)
// and writes z + k sZ (and vw + k sV).
struct FnoIO {
  const float* chi = nullptr;
  long sChi = 0;
  const float2* v = nullptr;
  long sV = 0;
  const double* nrm = nullptr;
  long sNrm = 0;
  float2* vw = nullptr;
  float2* z = nullptr;
  long sZ = 0;
};
int fno_forward_io(cgm_fno* f, hipStream_t st, const FnoIO& io, int B, bool train, const int* done);
int fno_batch(const cgm_fno* f);
cgm_context* fno_context(const cgm_fno* f);

// Tile preconditioner (cgm_tiles.hip): z = P_tiles(v / nrm) for nsys systems
// with the strides of io (chi unused: the tiles were factored for their maps);
// system s is skipped once *(done + s sd bytes).
int tiles_apply_io(const cgm_tiles* t, hipStream_t st, const FnoIO& io, int nsys, const int* done, long sd);
cgm_context* tiles_context(const cgm_tiles* t);
int tiles_maps(const cgm_tiles* t);

// The right preconditioner of the flexible solvers: the tile inverses, the
// network, or the identity (both null).
struct Precond {
  cgm_fno* net = nullptr;
  cgm_tiles* tiles = nullptr;
};
}  // namespace cgm

struct cgm_context {
  int N = 0, M = 0, device = 0;
  cgm::LrTiming lrt;   // live timing of the long-recurrence kernels (off by default)
  hipStream_t stream = nullptr;
  cgm::Plan plan{};    // complex64 passes (autotuned tiles)
  cgm::Plan plan64{};  // complex128 passes and the spectrum precompute (default tiles)
  float2* tw32 = nullptr;
  double2* tw64 = nullptr;
  float2* T32 = nullptr;   // N x M scratch of one complex64 matvec
  double2* T64 = nullptr;  // N x M scratch of one complex128 matvec (lazy)
  float mask_chi = 0.f;
  float tuned_us[3] = {0.f, 0.f, 0.f};  // autotuned pass 1/2/3 times
  float tuned_wide_us[3][3] = {}, tuned_lds_us[3][3] = {};  // [pass][batch 1/4/16], microseconds
  int path = 0;  // Arnoldi kernels: 0 = fused lagged CGS2 where supported, 1 = two-pass CGS2
  cgm::KrylovWs ws;
  cgm::LrWs lr;                  // long-recurrence FGMRES workspace (lazy)
  // FP64 matvecs of the long-recurrence steps (cgm_set_flexible_fp64)
  const double2* flex_g64 = nullptr;  // caller's FP64 spectrum (not owned)
  double* chi64 = nullptr;            // contrast maps in double (chi64_len values, lazy)
  double2* z64 = nullptr;             // z_j, x or b in complex128 (v64_len, lazy)
  double2* w64 = nullptr;             // w = A z_j or the residual in complex128 (v64_len, lazy)
  double2* T64B = nullptr;            // FP64 FFT scratch of the systems (T64B_n x N x M, lazy)
  long chi64_len = 0, v64_len = 0;
  int T64B_n = 0;
  cgm::DevState* st = nullptr;   // device
  cgm::DevState* hst = nullptr;  // pinned host mirror
  float* chi_tmp = nullptr;      // mask -> contrast conversions (N*N)
  float2* vec_tmp = nullptr;     // right-hand sides built on the device (N*N)
  double2* les_ws = nullptr;     // cgm_cuSolve_LES: FP64 copy of A (kMaxRestart^2, lazy)
  double2* v64a = nullptr;       // refinement residual (N*N, lazy)
  float* chi32 = nullptr;        // refinement contrast in float (N*N, lazy)
  float2* zb = nullptr;          // FGMRES preconditioned basis (zb_len elements, lazy)
  float2* pc_v = nullptr;        // two-level preconditioner: normalised inputs (2 pc_len, lazy)
  float2* pc_r = nullptr;        //   residuals of the tile solves (pc_len)
  float2* pc_z = nullptr;        //   network corrections (pc_len)
  long pc_len = 0;
  long zb_len = 0;
  cgm::DevState* stB = nullptr;   // batched solves: one state per system (device)
  cgm::DevState* hstB = nullptr;  // pinned host mirror
  int stB_n = 0;
  float2* T32B = nullptr;         // batched FFT scratch (T32B_n systems x N x M)
  int T32B_n = 0;
  hipStream_t gstream = nullptr;  // non-blocking stream of captured batched cycles (lazy)
  int use_graphs = 1;             // capture and replay batched restart cycles
};
