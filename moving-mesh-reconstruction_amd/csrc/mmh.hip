// mmh.hip -- host orchestration and C ABI of the MI355X moving-mesh reconstruction library.
//
// One context owns a HIP stream and a cache of device work buffers.  The V-cycle runs entirely
// on the device: residual norms, guard decisions and mean removals are device scalars consumed by
// later kernels, so a cycle issues no host synchronisation and no allocation once its buffers
// exist.  Declarations and reference citations of the entry points are in include/mmh.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mmh.h"
#include "kernels.hpp"
#include "kernels_lim.hpp"
#include "kernels_adj.hpp"
#include "kernels_wave.hpp"
#include "kernels_pd.hpp"
#include "kernels_vis.hpp"
#include "kernels_march.hpp"
#include "kernels_coarse.hpp"
#include "kernels_gal.hpp"

#include <hipcub/hipcub.hpp>

using namespace mmh;

struct mmh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::string, std::pair<void*, size_t>> ws;
  std::set<std::string> uploaded;  // constant tables already copied to the device
  std::string err;
  int sweep_impl = 0;  // mmh_time_sweeps layout: 0 sub-lattice (the V-cycle's), 1 column-major
  int mixed_delta = 0;     // FP32 preconditioner stencil in difference form (option "mixed_delta")
  int wave = 0;            // whole-sweep wavefront kernel for levels with n >= wave_min_n (opt-in)
  int wave_min_n = 128;
  int wave_p = 1;          // points per thread in flight per colour phase of the wavefront kernel
  int march_op = 1;        // residual/operator of large levels by k-marching workgroups (option "march_op")
  int march_tile = 0;      // tile shape of the k-marching kernels (option "march_tile": 0, 1, 2)
  int galerkin = 1;        // Krylov preconditioner: Galerkin coarse operators (option "galerkin")
  int nu_top = 0;          // Krylov preconditioner: sweeps on the finest level, 0 = nu_pre (option "nu_top")
  int coarse_fused = 16;   // Krylov preconditioner: levels <= this size in one workgroup (option "coarse_fused", 0 off)
  int sweep_slabs = 4;  // mixed preconditioner's finest level (n >= 512): colour sweeps slab by slab (option "sweep_slabs", 0 off)
  int slab_now = 0;     // set while such sweeps are issued
  int coarse_lds = 1;      // one-workgroup coarse cycle: its smallest levels held in LDS (option "coarse_lds")
  int dots32 = 1;          // mixed Krylov solve: z.r and z.r_old from the FP32 roundings of r (option "dots32")
  int coarse_fused_ref = 16;  // the reference V-cycle's levels <= this size in one workgroup (option "coarse_fused_ref")
  int zero_start = 1;      // first sweep from u = 0 reads only the colour classes already updated (option "zero_start")
  int colour_col = 8;      // colour kernels of levels n >= colour_col_min_n: points per thread along k (option "colour_col": 1 .. 16)
  int colour_col_min_n = 256;  // below 256^3 the n^3/(8S) threads are too few to fill the GPU
  int pcg_graph = 1;       // single solve: replay the PCG preconditioner as a captured HIP graph
  int batch_graph = 1;     // batched solve: replay one PCG iteration as a captured HIP graph
  int graphs_live = 0;     // captured graphs holding workspace pointers (workspace must not move)
  int batch_autotune = 1;  // batched solve: pick the colour kernel's launch shape by timing
  int batch_streams = 1;   // batched solve: size groups iterate side by side on streams of their own
  int batch_coarse_fused = 16;  // batched solve: levels <= this size one workgroup per grid (option "batch_coarse_fused", 0 off)
  int batch_persistent = 0;  // batched solve: groups of n <= this solve in one workgroup per grid (option "batch_persistent")
  int precond_bf16 = 1;    // mixed Krylov solve: 16-bit coefficients in the FP32 preconditioner (option "precond_bf16")
  int mesh_solve = 0;      // calcdefp's mesh solve: 0 one reference V-cycle, 1 V(2,2) cycles to 10^-mesh_tol,
                           // 2 the Krylov solve to 10^-mesh_tol (option "mesh_solve")
  int mesh_tol = 8;        // option "mesh_tol": the tolerance exponent of mesh_solve 1 and 2
  int mesh_cycles = 0;     // cycles (or iterations) of the last calcdefp mesh solve
  std::map<std::string, std::pair<int, int>> tune;  // (grid size, batch) -> (block size, XCD remap)
  struct Arena* arena = nullptr;  // set while an allocation-free entry point runs: ws_get carves from it
};

// Workspace of the allocation-free entry points: buffers carved in request order from the
// device's static scratch array; nothing outlives the call.
struct Arena {
  unsigned char* base = nullptr;
  size_t cap = 0, off = 0;
  std::map<std::string, std::pair<void*, size_t>> ws;
};

// Scratch memory of the allocation-free entry point (mmh_multigrid_noalloc_d): one array per GPU,
// placed by the code object loader and held until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, and the device address is
// taken once per device with hipGetSymbolAddress (scratch_setup) and used instead.
#define MMH_SCRATCH_BYTES (((size_t)32) << 20)
static __device__ __align__(256) unsigned char scratch[MMH_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;    // device -> address of that device's scratch array
std::map<int, hipError_t> g_scratch_rc;  // device -> result of the address lookup

// Called by mmh_create with `device` current: the first context on a device looks the address up.
void scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_rc.count(device)) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_rc[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = ptr;
}

void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  return it == g_scratch_ptr.end() ? nullptr : it->second;
}

struct MmhError : std::runtime_error {
  int code;
  MmhError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess)                                                                              \
      throw MmhError(e_ == hipErrorOutOfMemory ? MMH_ENOMEM : MMH_EDEVICE,                             \
                     std::string(#x) + ": " + hipGetErrorString(e_));                                  \
  } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

inline void require(bool c, const char* msg) {
  if (!c) throw MmhError(MMH_EINVAL, msg);
}
inline bool is_pow2(int n) { return n >= 4 && n <= 4096 && (n & (n - 1)) == 0; }
__host__ __device__ inline i64 cube(int n) { return (i64)n * n * n; }
// Blocks for n work-items.  A dispatch carries its grid size in work-items in 32 bits, so larger
// launches are split by the caller; reaching this check means one was not.
inline unsigned nblk(i64 n, int b = 256) {
  if (n > (i64)0xFFFFFFFFLL - b) throw MmhError(MMH_EINVAL, "internal: launch of more than 2^32 work-items");
  return (unsigned)((n + b - 1) / b);
}

void* ws_get(mmh_ctx* c, const std::string& key, size_t bytes) {
  if (c->arena) {
    Arena& A = *c->arena;
    auto ai = A.ws.find(key);
    if (ai != A.ws.end() && ai->second.second >= bytes) return ai->second.first;
    const size_t b = (std::max<size_t>(bytes, 256) + 255) / 256 * 256;
    if (A.off + b > A.cap)
      throw MmhError(MMH_EINVAL, "allocation-free entry: the problem needs more than the " +
                                     std::to_string(A.cap >> 20) + " MiB scratch array");
    void* p = A.base + A.off;
    A.off += b;
    A.ws[key] = {p, bytes};
    return p;
  }
  auto it = c->ws.find(key);
  if (it != c->ws.end() && it->second.second >= bytes) return it->second.first;
  if (it != c->ws.end()) {
    // a captured graph keeps the old address in its kernel arguments
    if (c->graphs_live) throw MmhError(MMH_EDEVICE, "internal: workspace '" + key + "' grows while a graph is live");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(it->second.first));
    c->ws.erase(it);
  }
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  c->ws[key] = {p, bytes};
  return p;
}
// Frees one cached buffer (after the queued work that may use it).
void ws_free(mmh_ctx* c, const std::string& key) {
  if (c->arena) return;  // the arena is dropped as a whole when the call ends
  auto it = c->ws.find(key);
  if (it == c->ws.end()) return;
  if (c->graphs_live) throw MmhError(MMH_EDEVICE, "internal: workspace '" + key + "' freed while a graph is live");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(it->second.first));
  c->ws.erase(it);
}
// Bytes a wsT request would newly allocate (0 if the key already holds enough).
size_t ws_need(const mmh_ctx* c, const std::string& key, size_t bytes) {
  auto it = c->ws.find(key);
  return (it != c->ws.end() && it->second.second >= bytes) ? 0 : bytes;
}

template <class T> T* wsT(mmh_ctx* c, const std::string& key, i64 count) {
  return (T*)ws_get(c, key + ":" + std::to_string(sizeof(T)), sizeof(T) * (size_t)count);
}

template <class T> void fill(mmh_ctx* c, T* a, T v, i64 n) {
  constexpr i64 CH = (i64)1 << 30;
  for (i64 o = 0; o < n; o += CH) {
    const i64 m = std::min(CH, n - o);
    k_fill<T><<<nblk(m), 256, 0, c->stream>>>(a + o, v, m);
  }
  LAUNCHCHK();
}

// Graph capture needs a stream other than the null stream: work that is captured runs on a stream
// of the library's own (default priority), ordered after the work already queued on the context's
// stream; the context's stream is restored, after a synchronisation, when the guard goes.
struct StreamSwap {
  mmh_ctx* c;
  hipStream_t saved, own = nullptr;
  hipEvent_t ev = nullptr;
  explicit StreamSwap(mmh_ctx* cc) : c(cc), saved(cc->stream) {
    HIPCHK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, saved));
    HIPCHK(hipStreamWaitEvent(own, ev, 0));
    c->stream = own;
  }
  ~StreamSwap() {
    (void)hipStreamSynchronize(own);
    c->stream = saved;
    (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(own);
  }
};

// A sequence of launches captured once on the context's current stream and then replayed on that
// stream.  While it exists, workspace buffers may be added but not moved (ws_get refuses to grow
// one).
struct Replay {
  mmh_ctx* c;
  hipStream_t st = nullptr;  // the stream it was captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  explicit Replay(mmh_ctx* cc) : c(cc) {}
  template <class F> void capture(F&& work) {
    st = c->stream;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      work();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &graph));
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    ++c->graphs_live;
  }
  void launch() { HIPCHK(hipGraphLaunch(exec, st)); }
  ~Replay() {
    if (exec) {
      (void)hipStreamSynchronize(st);
      (void)hipGraphExecDestroy(exec);
      --c->graphs_live;
    }
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// FP64 sum (op 0) or max|x| (op 1) of x[0..n) into the device double *dst, blocked tree order.
template <class T> void reduce(mmh_ctx* c, const T* x, i64 n, int op, double* dst) {
  const i64 ch = (i64)RB * RPT;
  i64 nb = (n + ch - 1) / ch;
  if (nb <= 1) {
    if (op == 0) k_reduce_stage<T, 0><<<1, RB, 0, c->stream>>>(dst, x, n);
    else k_reduce_stage<T, 1><<<1, RB, 0, c->stream>>>(dst, x, n);
    LAUNCHCHK();
    return;
  }
  double* a = (double*)ws_get(c, "red_a", sizeof(double) * nb);
  double* b = (double*)ws_get(c, "red_b", sizeof(double) * ((nb + ch - 1) / ch + 1));
  if (op == 0) k_reduce_stage<T, 0><<<(unsigned)nb, RB, 0, c->stream>>>(a, x, n);
  else k_reduce_stage<T, 1><<<(unsigned)nb, RB, 0, c->stream>>>(a, x, n);
  LAUNCHCHK();
  while (nb > 1) {
    i64 nb2 = (nb + ch - 1) / ch;
    double* out = nb2 == 1 ? dst : b;
    if (op == 0) k_reduce_stage<double, 0><<<(unsigned)nb2, RB, 0, c->stream>>>(out, a, nb);
    else k_reduce_stage<double, 1><<<(unsigned)nb2, RB, 0, c->stream>>>(out, a, nb);
    LAUNCHCHK();
    std::swap(a, b);
    nb = nb2;
  }
}

// ------------------------------------------------------------------------------------------------
// Multigrid hierarchy.  Level 0 aliases the caller's arrays; coarser levels live in the context.
template <class T> struct Hier {
  int nlev = 0;
  int n[MMH_MAXLEV];
  int nu[MMH_MAXLEV];
  T dx[MMH_MAXLEV];
  T* u[MMH_MAXLEV];
  T* rhs[MMH_MAXLEV];
  const T* def[MMH_MAXLEV];
  const T* u1[MMH_MAXLEV];
  T* G[MMH_MAXLEV];
  T* r[MMH_MAXLEV];
  T* A[MMH_MAXLEV] = {};  // Galerkin coarse operators (27 fields per level), null where not built
  // 16-bit copies of G / A that the FP32 preconditioner reads instead, where set (option
  // "precond_bf16"); the arithmetic stays FP32
  const bf16* Gb[MMH_MAXLEV] = {};
  const bf16x2* Ab[MMH_MAXLEV] = {};  // 16-bit Galerkin coefficients, two per word
  double* scal;  // per level: [0] r1, [1] relaxerr, [2] sum; then [3*MAXLEV] rhs sums
  int* guard;    // per level, plus entry flag at [MAXLEV]
};

struct Sched {
  int kind = 0;  // 0 reference, 1 fixed
  int ng = 0;
  int nu_pre = 2, nu_post = 2, nu_coarse = 32;
  bool guard = true;       // the reference's divergence guard on coarse levels
  bool symmetric = false;  // post-smoothing in reverse colour order (symmetric V-cycle)
  bool delta = false;      // difference form of the FP32 stencil (mixed preconditioner only)
  bool defer_top_mean = false;  // finest level: sum only, its consumers remove the mean on the fly
  bool sum_by_dots = false;     // with defer_top_mean: not even the sum (the dot-product pass forms it)
  int nu_top = 0;  // kind 1: sweeps each way on the finest level when > 0 (option "nu_top")
  bool top_zero = false;  // the finest level's u is 0 on entry (a preconditioner's z); coarser ones always are
  int pre(int n) const { return kind == 0 ? 4 * ng / n : (nu_top > 0 && n == ng ? nu_top : nu_pre); }
  int post(int n) const { return kind == 0 ? 4 * ng / n : (nu_top > 0 && n == ng ? nu_top : nu_post); }
  int coarse() const { return kind == 0 ? 32 : nu_coarse; }
};

// Builds level arrays and metrics for a problem of size n with defpot/ubig at level 0.
template <class T>
void build_hier(mmh_ctx* c, Hier<T>& H, const std::string& tag, T* u, T* rhs, const T* defpot, const T* ubig, int nu, T dx,
                int n, int lsquare, bool want_G0 = true, T* G0 = nullptr) {
  int nl = 0;
  for (int m = n; m >= 4; m >>= 1) ++nl;
  require(nl <= MMH_MAXLEV, "too many levels");
  H.nlev = nl;
  H.scal = (double*)ws_get(c, tag + "scal", sizeof(double) * 4 * MMH_MAXLEV);
  H.guard = (int*)ws_get(c, tag + "guard", sizeof(int) * (MMH_MAXLEV + 1));
  for (int l = 0; l < nl; ++l) {
    const int m = n >> l;
    H.n[l] = m;
    H.dx[l] = l == 0 ? dx : H.dx[l - 1] * T(2);
    const std::string t = tag + std::to_string(l);
    if (l == 0) {
      H.u[0] = u; H.rhs[0] = rhs; H.def[0] = defpot; H.u1[0] = ubig; H.nu[0] = nu;
      if (nu > 1) {
        // density component of the AoS fluid array, made contiguous once for the metric's box mean
        T* w = wsT<T>(c, t + "u1", cube(m));
        k_copy_strided<T><<<nblk(cube(m)), 256, 0, c->stream>>>(w, ubig, cube(m), nu);
        LAUNCHCHK();
        H.u1[0] = w; H.nu[0] = 1;
      }
    } else {
      H.u[l] = wsT<T>(c, t + "u", cube(m));
      H.rhs[l] = wsT<T>(c, t + "rhs", cube(m));
      T* d = wsT<T>(c, t + "def", cube(m));
      T* w = wsT<T>(c, t + "u1", cube(m));
      const int pm = n >> (l - 1);
      k_restrict27<T><<<nblk(cube(m)), 256, 0, c->stream>>>(d, H.def[l - 1], pm, 1);
      k_restrict27<T><<<nblk(cube(m)), 256, 0, c->stream>>>(w, H.u1[l - 1], pm, H.nu[l - 1]);
      LAUNCHCHK();
      H.def[l] = d; H.u1[l] = w; H.nu[l] = 1;
    }
    H.r[l] = wsT<T>(c, t + "r", cube(m));
    if (l == 0 && G0) H.G[0] = G0;
    else H.G[l] = wsT<T>(c, t + "G", 6 * cube(m));
    if (l > 0 || want_G0) {
      k_metric<T, 1><<<nblk(cube(m)), 256, 0, c->stream>>>(H.G[l], H.def[l], H.u1[l], H.nu[l], m, H.dx[l], lsquare);
      LAUNCHCHK();
    }
  }
}

// The Galerkin operators take 27 values per coarse point (about 3.9 fine-grid words, plus half
// that again in FP32 for the mixed preconditioner).  They are built only where they fit, with the
// rest of the Krylov solve's workspace (its vectors, the FP32 hierarchy) and 4 GiB to spare, beside
// the memory already held; otherwise the rebuilt coarse operators are used.
// The FP32 copies counted are those build_fp32 makes (bf: 16-bit coefficients on levels >= 32^3).
// d32: the mixed Krylov solve's previous residual is kept in FP32 (option "dots32").
bool galerkin_fits(mmh_ctx* c, const Hier<double>& H, const std::string& tag, bool mixed, bool bf, bool d32) {
  size_t need = 0;
  const size_t n3 = (size_t)cube(H.n[0]);
  for (const char* k : {"cg_r:8", "cg_p:8", "cg_q:8", "det:8"}) need += ws_need(c, k, n3 * sizeof(double));
  need += d32 ? ws_need(c, "cg_ro32:4", n3 * sizeof(float)) : ws_need(c, "cg_ro:8", n3 * sizeof(double));
  if (!mixed) need += ws_need(c, "cg_z:8", n3 * sizeof(double));
  for (int l = 0; l < H.nlev; ++l) {
    const size_t m3 = (size_t)cube(H.n[l]);
    const std::string t = "sf" + std::to_string(l);
    if (l > 0) need += ws_need(c, tag + std::to_string(l) + "A:8", 27 * m3 * sizeof(double));
    if (!mixed) continue;
    const bool b16 = bf && H.n[l] >= 32;
    if (l > 0) need += b16 ? ws_need(c, t + "Ab2", 14 * m3 * sizeof(bf16x2)) : ws_need(c, t + "A:4", 27 * m3 * sizeof(float));
    if (b16 && l == 0) need += ws_need(c, t + "Gb", 6 * m3 * sizeof(bf16));
    else if (!b16) need += ws_need(c, t + "G:4", 6 * m3 * sizeof(float));
    need += ws_need(c, t + "u:4", m3 * sizeof(float)) + ws_need(c, t + "rhs:4", m3 * sizeof(float)) +
            ws_need(c, t + "r:4", m3 * sizeof(float));
  }
  size_t freeb = 0, total = 0;
  HIPCHK(hipMemGetInfo(&freeb, &total));
  return need + ((size_t)4 << 30) <= freeb;
}

// Galerkin operators A_l = R A_{l-1} P of levels 1.. (FP64), the first from the fine metric.
void build_galerkin(mmh_ctx* c, Hier<double>& H, const std::string& tag) {
  for (int l = 1; l < H.nlev; ++l) {
    const int m = H.n[l];
    H.A[l] = wsT<double>(c, tag + std::to_string(l) + "A", 27 * cube(m));
    if (l == 1)
      k_rap19<double><<<nblk(cube(m)), 256, 0, c->stream>>>(H.A[1], H.G[0], H.n[0], 1.0 / (H.dx[0] * H.dx[0]));
    else
      k_rap27<<<nblk(cube(m)), 256, 0, c->stream>>>(H.A[l], H.A[l - 1], H.n[l - 1]);
    LAUNCHCHK();
  }
}

// Grid of the full-field stencil kernels (k_resid, k_apply_dot): one thread per point in layout 0,
// one per 2x2x2 cell in layout 1.
inline unsigned stencil_grid(int n, int layout) { return layout == 1 ? nblk(cube(n / 2), RB) : nblk(cube(n), RB); }

// Tiles (TI x TJ points) of the k-marching kernels, by option "march_tile".  0 (default): the
// fastest measured at 512^3 on MI355X, 64 x 8 (512 threads) in FP64 and 64 x 4 (256 threads) in
// FP32; 1 and 2 are the alternatives that were timed against them.
template <class T, int V> struct MarchTile;
template <> struct MarchTile<double, 0> { static constexpr int TI = 64, TJ = 8; };
template <> struct MarchTile<double, 1> { static constexpr int TI = 32, TJ = 8; };
template <> struct MarchTile<double, 2> { static constexpr int TI = 32, TJ = 16; };
template <> struct MarchTile<float, 0> { static constexpr int TI = 64, TJ = 4; };
template <> struct MarchTile<float, 1> { static constexpr int TI = 64, TJ = 8; };
template <> struct MarchTile<float, 2> { static constexpr int TI = 128, TJ = 4; };

// Planes per workgroup of k_op_march for a level of size n: the largest of 32, 16, 8 leaving at
// least 2048 workgroups (8 at least, so the block results never outnumber stencil_grid's); 0 where
// the marching kernel does not apply (small levels, the option off).
template <class T, int V> int march_ks(const mmh_ctx* c, int n) {
  constexpr int TI = MarchTile<T, V>::TI, TJ = MarchTile<T, V>::TJ;
  if (!c->march_op || n < 64 || n > 1024 || n % TI || n % TJ || n % 8) return 0;  // 32-bit offsets up to 1024^3
  for (int ks = 32; ks > 8; ks >>= 1)
    if (n % ks == 0 && cube(n) / ((i64)TI * TJ * ks) >= 2048) return ks;
  return 8;
}

// One k_op_march launch (MODE 0 residual, 1 operator) if the level qualifies; returns the number
// of block results, 0 when it does not apply.
template <class T, int MODE, int V, class TG = T>
unsigned march_launch(mmh_ctx* c, T* out, double* bres, const T* phi, const T* rhs, const TG* G, int n, T di2, int B,
                      int batched, const int* act) {
  constexpr int TI = MarchTile<T, V>::TI, TJ = MarchTile<T, V>::TJ;
  const int ks = march_ks<T, V>(c, n);
  if (!ks) return 0;
  const unsigned nb = (unsigned)(cube(n) / ((i64)TI * TJ * ks));
  k_op_march<T, MODE, TI, TJ, TG><<<dim3(nb, B), TI * TJ, 0, c->stream>>>(out, bres, phi, rhs, G, n, di2, ks, batched,
                                                                          act);
  LAUNCHCHK();
  return nb;
}
template <class T, int MODE, class TG = T>
unsigned march_op(mmh_ctx* c, T* out, double* bres, const T* phi, const T* rhs, const TG* G, int n, T di2, int B,
                  int batched, const int* act = nullptr) {
  switch (c->march_tile) {
    case 1: return march_launch<T, MODE, 1, TG>(c, out, bres, phi, rhs, G, n, di2, B, batched, act);
    case 2: return march_launch<T, MODE, 2, TG>(c, out, bres, phi, rhs, G, n, di2, B, batched, act);
    default: return march_launch<T, MODE, 0, TG>(c, out, bres, phi, rhs, G, n, di2, B, batched, act);
  }
}

// r = rhs - L phi on a sub-lattice level (r may be null) with per-block max|r| into bres, for B
// grids stored one after another when batched (act: grids y with act[y] = 0 are skipped); returns
// the number of block results per grid.
template <class T, class TG = T>
unsigned op_resid(mmh_ctx* c, T* r, double* bres, const T* phi, const T* rhs, const TG* G, int n, T di2,
                  bool delta = false, int B = 1, int batched = 0, const int* act = nullptr) {
  if (!delta)
    if (const unsigned nb = march_op<T, 0, TG>(c, r, bres, phi, rhs, G, n, di2, B, batched, act)) return nb;
  const unsigned nb = stencil_grid(n, 1);
  if (delta) k_resid<T, 1, true, TG><<<dim3(nb, B), RB, 0, c->stream>>>(r, bres, phi, rhs, G, n, di2, batched, act);
  else k_resid<T, 1, false, TG><<<dim3(nb, B), RB, 0, c->stream>>>(r, bres, phi, rhs, G, n, di2, batched, act);
  LAUNCHCHK();
  return nb;
}

// q = L p with per-block sums of p*q into bres (same conventions as op_resid).
template <class T>
unsigned op_apply(mmh_ctx* c, T* q, double* bres, const T* p, const T* G, int n, T di2, int B = 1, int batched = 0,
                  const int* act = nullptr) {
  if (const unsigned nb = march_op<T, 1>(c, q, bres, p, (const T*)nullptr, G, n, di2, B, batched, act)) return nb;
  const unsigned nb = stencil_grid(n, 1);
  k_apply_dot<T, 1><<<dim3(nb, B), RB, 0, c->stream>>>(q, bres, p, G, n, di2, batched, act);
  LAUNCHCHK();
  return nb;
}

template <class T> void relayout(mmh_ctx* c, T* dst, const T* src, int n, int comps, int to_sl) {
  const i64 cnt = cube(n) * comps;
  (void)cnt;
  const i64 n3 = cube(n);
  for (int f = 0; f < comps; ++f) {  // one field per launch (n^3 <= 2^30 work-items)
    if (to_sl) k_relayout<T, 1><<<nblk(n3), 256, 0, c->stream>>>(dst + f * n3, src + f * n3, n, 1);
    else k_relayout<T, 0><<<nblk(n3), 256, 0, c->stream>>>(dst + f * n3, src + f * n3, n, 1);
  }
  LAUNCHCHK();
}

// ns colour sweeps of phi: 8 launches per sweep, one per colour class; layout 1 (sub-lattice, the
// V-cycle's) or 0 (column-major).
// Tile shape of the wavefront sweep: FP32 64 x 32, FP64 32 x 32 points (four LDS planes stay
// under 64 KiB, two workgroups per CU).
template <class T> struct WaveTile;
template <> struct WaveTile<float> { static constexpr int TI = 64, TJ = 32; };
template <> struct WaveTile<double> { static constexpr int TI = 32, TJ = 32; };

template <class T> bool wave_fits(mmh_ctx* c, int n) {
  constexpr int TI = WaveTile<T>::TI, TJ = WaveTile<T>::TJ;
  return c->wave && n >= c->wave_min_n && n >= TI + 2 * WAVE_HI && n >= TJ + 2 * WAVE_HJ && n % TI == 0 && n % TJ == 0;
}

// One wavefront sweep from `in` into `out` (forward or reverse colour order).
template <class T> void wave_sweep(mmh_ctx* c, T* out, const T* in, const T* rhs, const T* G, int n, T dx2, bool rev) {
  constexpr int TI = WaveTile<T>::TI, TJ = WaveTile<T>::TJ;
  const i64 tiles = (i64)(n / TI) * (n / TJ);
  int K = 8;  // planes of each parity per segment: the largest of 8, 4, 2 giving >= 1024 workgroups
  while (K > 2 && tiles * (n / (2 * K)) < 1024) K >>= 1;
  require(n % (2 * K) == 0, "wave sweep: n must be a multiple of 2K");
  const unsigned nwg = (unsigned)(tiles * (n / (2 * K)));
  const size_t lds = sizeof(T) * 4 * (TI + 2 * WAVE_HI) * (TJ + 2 * WAVE_HJ);
#define WAVE_LAUNCH(P)                                                                                         \
  if (rev) k_sweep_wave<T, TI, TJ, true, P><<<nwg, 512, lds, c->stream>>>(out, in, rhs, G, n, dx2, K);        \
  else k_sweep_wave<T, TI, TJ, false, P><<<nwg, 512, lds, c->stream>>>(out, in, rhs, G, n, dx2, K);
  if (c->wave_p == 4) { WAVE_LAUNCH(4) }
  else if (c->wave_p == 2) { WAVE_LAUNCH(2) }
  else { WAVE_LAUNCH(1) }
#undef WAVE_LAUNCH
  LAUNCHCHK();
}

// One colour class by k_colour_col: S points per thread along k (the same values as k_colour).
template <class T, class TG = T>
void colour_col_launch(mmh_ctx* c, int S, bool masked, bool delta, T* phi, const T* rhs, const TG* G, int n, T dx2,
                       int kp, int io, int jo, unsigned live) {
  const int h = n / 2;
  require(S > 1 && h % S == 0, "colour_col: points per thread must divide n/2");
  const unsigned nb = colour_col_blocks(n) * (unsigned)(h / S);
#define COL_LAUNCH(SS)                                                                                           \
  if (masked) k_colour_col<T, false, true, SS, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo, live); \
  else if (delta) k_colour_col<T, true, false, SS, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo);   \
  else k_colour_col<T, false, false, SS, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo);
  if (S == 2) { COL_LAUNCH(2) }
  else if (S == 4) { COL_LAUNCH(4) }
  else if (S == 8) { COL_LAUNCH(8) }
  else { COL_LAUNCH(16) }
#undef COL_LAUNCH
}

// from_zero: phi is 0 on entry, so the first sweep's colours read only the classes updated before
// them in that sweep (k_colour's MASK form; same values).
template <class T, class TG = T>
void sweeps(mmh_ctx* c, T* phi, const T* rhs, const TG* G, int n, T dx, int ns, int layout = 1, bool reverse = false,
            bool delta = false, bool from_zero = false) {
  const T dx2 = dx * dx;
  const int h = n / 2;
  const unsigned nb = nblk(cube(h));
  from_zero = from_zero && layout == 1 && !delta;
  if constexpr (std::is_same<T, TG>::value)
  if (layout == 1 && ns >= 2 && !delta && !from_zero && wave_fits<T>(c, n)) {
    // pairs of sweeps through a second buffer; an odd last sweep goes colour by colour below
    T* tmp = wsT<T>(c, "wave_tmp" + std::to_string(n), cube(n));
    const int pairs = ns / 2;
    for (int q = 0; q < pairs; ++q) {
      wave_sweep<T>(c, tmp, phi, rhs, G, n, dx2, reverse);
      wave_sweep<T>(c, phi, tmp, rhs, G, n, dx2, reverse);
    }
    ns -= 2 * pairs;
  }
  const int S = (layout == 1 && n >= c->colour_col_min_n && h % c->colour_col == 0) ? c->colour_col : 1;
  if (c->slab_now > 1 && S == 8 && !delta && n >= 512 && (h / S) % c->slab_now == 0) {
    // Slab order (the mixed preconditioner's finest level): all 8 colours on one k-slab of the
    // level, then the next slab; reversed, the exact reverse sequence, so the symmetric V-cycle stays
    // symmetric.  A slab's class arrays and coefficients are a fraction of the level, and the
    // colour launches after the first find part of them in the 256 MB Infinity Cache.  A different
    // Gauss-Seidel order from the colour-by-colour sweep (this is a preconditioner, not the
    // reference's relaxation); phi must hold its real values (0 for a sweep from zero).
    const int slabs = c->slab_now, per = h / S / slabs;
    const unsigned nbs = colour_col_blocks(n) * (unsigned)per;
    for (int s = 0; s < ns; ++s)
      for (int bi = 0; bi < slabs; ++bi) {
        const int b = reverse ? slabs - 1 - bi : bi;
        for (int cc = 0; cc < 8; ++cc) {
          const int col = reverse ? 7 - cc : cc;
          k_colour_col<T, false, false, 8, TG><<<nbs, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, col >> 2, (col >> 1) & 1,
                                                                        col & 1, 0xFFu, b * per, per);
        }
      }
    LAUNCHCHK();
    return;
  }
  for (int s = 0; s < ns; ++s)
    for (int cc = 0, done = 0; cc < 8; ++cc) {
      const int col = reverse ? 7 - cc : cc;
      const int kp = col >> 2, io = (col >> 1) & 1, jo = col & 1;
      if (S > 1)
        colour_col_launch<T, TG>(c, S, s == 0 && from_zero, delta, phi, rhs, G, n, dx2, kp, io, jo, (unsigned)done);
      else if (s == 0 && from_zero)
        k_colour<T, 1, false, true, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo, 0, 1,
                                                                   (unsigned)done);
      else if (layout == 1 && delta)
        k_colour<T, 1, true, false, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo);
      else if (layout == 1) k_colour<T, 1, false, false, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo);
      else k_colour<T, 0, false, false, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, G, n, dx2, kp, io, jo);
      done |= 1 << col;
    }
  LAUNCHCHK();
}

// relax on the device (level arrays in sub-lattice layout): sweeps, then residual (stored when
// rc != null), max-norm into *rinf_dev, and the calcresid restriction into rc.
template <class T, class TG = T>
void relax_dev(mmh_ctx* c, T* phi, const T* rhs, const TG* G, T* rscratch, int n, T dx, int nrelax, int want_norm, T* rc,
               double* rinf_dev, bool reverse = false, bool delta = false, bool from_zero = false) {
  sweeps<T, TG>(c, phi, rhs, G, n, dx, nrelax, 1, reverse, delta, from_zero);
  if (!want_norm && !rc) return;
  double* bmax = (double*)ws_get(c, "bmax", sizeof(double) * stencil_grid(n, 1));
  const unsigned nb = op_resid<T, TG>(c, rc ? rscratch : nullptr, bmax, phi, rhs, G, n, T(1) / (dx * dx), delta);
  if (want_norm) reduce<double>(c, bmax, nb, 1, rinf_dev);
  if (rc) {
    const int nrmax = std::min(4, std::max(1, n / 4));
    int nrl = nrelax > nrmax ? (nrelax % nrmax ? nrelax % nrmax : nrmax) : nrelax;
    k_restrict_resid<T, 1><<<nblk(cube(n / 2)), 256, 0, c->stream>>>(rc, rscratch, n, nrl % (n / 2));
    LAUNCHCHK();
  }
}

// Sweeps and relax of a Galerkin (27-point) level: the colour order of the 19-point sweeps, the
// residual stored in r and restricted into rc with calcresid's plane order rule.
template <class T, class TG = T>
void gal_sweeps(mmh_ctx* c, T* phi, const T* rhs, const TG* A, int n, int ns, bool reverse, bool from_zero = false) {
  const unsigned nb = nblk(cube(n / 2));
  for (int s = 0; s < ns; ++s)
    for (int cc = 0, done = 0; cc < 8; ++cc) {
      const int col = reverse ? 7 - cc : cc;
      if (s == 0 && from_zero)
        k_gal_colour<T, true, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, A, n, col >> 2, (col >> 1) & 1, col & 1, 0,
                                                             (unsigned)done);
      else k_gal_colour<T, false, TG><<<nb, 256, 0, c->stream>>>(phi, rhs, A, n, col >> 2, (col >> 1) & 1, col & 1);
      done |= 1 << col;
    }
  LAUNCHCHK();
}
template <class T, class TG = T>
void gal_relax(mmh_ctx* c, T* phi, const T* rhs, const TG* A, T* r, int n, int nrelax, T* rc, bool from_zero = false) {
  gal_sweeps<T, TG>(c, phi, rhs, A, n, nrelax, false, from_zero);
  k_gal_resid<T, TG><<<gal_resid_blocks(n), 256, 0, c->stream>>>(r, phi, rhs, A, n);
  const int nrmax = std::min(4, std::max(1, n / 4));
  const int nrl = nrelax > nrmax ? (nrelax % nrmax ? nrelax % nrmax : nrmax) : nrelax;
  k_restrict_resid<T, 1><<<nblk(cube(n / 2)), 256, 0, c->stream>>>(rc, r, n, nrl % (n / 2));
  LAUNCHCHK();
}

// The levels l.. of a hierarchy as one coarse cycle (k_coarse_cycle), when the schedule allows it:
// the Krylov preconditioner's fixed symmetric schedule without divergence guard, levels up to the
// size set by option "coarse_fused" (16 by default: with 32-bit indexing in the workgroup and the
// smallest levels in LDS, one workgroup beats the launches up to 16^3 with the 27-point Galerkin
// operators too; at 32^3 a level's sweeps move megabytes through one CU and the launches win).
// The reference schedule (kind 0: guard, nu = 4 NG/N) fuses its 19-point levels up to the size set
// by option "coarse_fused_ref" (16 by default: the 16^3..4^3 levels run 128, 256 and 32+32 sweeps,
// some 3000 launches per V-cycle otherwise; at 512^3 both take 88-89 ms per cycle, and fusing 32^3
// as well is slower, one CU then carrying the 32^3 level).
template <class T> bool coarse_cycle_applies(const mmh_ctx* c, const Hier<T>& H, int l, const Sched& S) {
  if (S.delta || H.nlev - l > CLEV_MAX) return false;
  for (int q = l; q < H.nlev; ++q)
    if (H.Gb[q] || H.Ab[q]) return false;  // the one-workgroup cycle reads FP32 coefficients
  if (S.kind == 0) return c->coarse_fused_ref && !H.A[l] && H.n[l] <= std::min(c->coarse_fused_ref, CTOP_MAX);
  return c->coarse_fused && !S.guard && S.kind == 1 && S.nu_pre == S.nu_post && !(S.nu_top > 0 && H.n[l] == S.ng) &&
         H.n[l] <= std::min(c->coarse_fused, CTOP_MAX);
}
template <class T> CoarseCycle<T> coarse_cycle_of(Hier<T>& H, int l, const Sched& S, int ng_top) {
  CoarseCycle<T> C{};
  C.nl = H.nlev - l;
  for (int q = 0; q < C.nl; ++q) {
    C.u[q] = H.u[l + q];
    C.rhs[q] = H.rhs[l + q];
    C.r[q] = H.r[l + q];
    C.G[q] = H.G[l + q];
    C.scal[q] = H.scal + 3 * (l + q);
    C.guard[q] = H.guard + l + q;
    C.dx[q] = H.dx[l + q];
    C.n[q] = H.n[l + q];
    C.A[q] = H.A[l + q];
    C.sw[q] = S.pre(H.n[l + q]);
  }
  C.coarse = S.coarse();
  C.ref = S.guard ? 1 : 0;
  C.top = H.n[l] == ng_top ? 1 : 0;
  C.symmetric = S.symmetric ? 1 : 0;
  return C;
}

// One V-cycle from level l (mgtemplate.f90 / mg*.f90; mg4 at multigrid.f90:94-121).
template <class T> void vcycle(mmh_ctx* c, Hier<T>& H, int l, int ng_top, const Sched& S) {
  const int n = H.n[l];
  const i64 n3 = cube(n);
  double* sc = H.scal + 3 * l;
  if (coarse_cycle_applies(c, H, l, S)) {
    CoarseCycle<T> C = coarse_cycle_of(H, l, S, ng_top);
    if (c->coarse_lds) {
      stage_levels(C);
      k_coarse_cycle_lds<T><<<1, CW, 0, c->stream>>>(C);
    } else {
      k_coarse_cycle<T><<<1, CW, 0, c->stream>>>(C);
    }
    LAUNCHCHK();
    return;
  }
  const bool zero = c->zero_start && (l > 0 || S.top_zero);
  if (n == 4) {
    const int nc = S.coarse();
    if (H.A[l]) {
      gal_sweeps<T>(c, H.u[l], H.rhs[l], H.A[l], 4, nc, false, zero);
      gal_sweeps<T>(c, H.u[l], H.rhs[l], H.A[l], 4, nc, S.symmetric);
    } else {
      relax_dev(c, H.u[l], H.rhs[l], H.G[l], H.r[l], 4, H.dx[l], nc, S.guard, (T*)nullptr, sc + 0, false, false, zero);
      relax_dev(c, H.u[l], H.rhs[l], H.G[l], H.r[l], 4, H.dx[l], nc, S.guard, (T*)nullptr, sc + 1, S.symmetric);
    }
    reduce<T>(c, H.u[l], n3, 0, sc + 2);
    k_zerosum_guard<T><<<1, 256, 0, c->stream>>>(H.u[l], sc + 2, sc, n3, (n == ng_top || !S.guard) ? 0 : 2, H.guard + l);
    LAUNCHCHK();
    return;
  }
  if (H.A[l]) gal_relax<T>(c, H.u[l], H.rhs[l], H.A[l], H.r[l], n, S.pre(n), H.rhs[l + 1], zero);
  else relax_dev(c, H.u[l], H.rhs[l], H.G[l], H.r[l], n, H.dx[l], S.pre(n), S.guard, H.rhs[l + 1], sc + 0, false, false,
                 zero);
  fill<T>(c, H.u[l + 1], T(0), cube(n / 2));
  vcycle(c, H, l + 1, ng_top, S);
  k_prolong<T, 1><<<nblk(cube(n / 2)), 256, 0, c->stream>>>(H.u[l], H.u[l + 1], n, 1);
  LAUNCHCHK();
  if (H.A[l]) gal_sweeps<T>(c, H.u[l], H.rhs[l], H.A[l], n, S.post(n), S.symmetric);
  else relax_dev(c, H.u[l], H.rhs[l], H.G[l], H.r[l], n, H.dx[l], S.post(n), S.guard, (T*)nullptr, sc + 1, S.symmetric);
  reduce<T>(c, H.u[l], n3, 0, sc + 2);
  if (S.defer_top_mean && l == 0) return;
  k_zerosum_guard<T><<<nblk(n3), 256, 0, c->stream>>>(H.u[l], sc + 2, sc, n3, (n == ng_top || !S.guard) ? 0 : 1,
                                                      H.guard + l);
  LAUNCHCHK();
}

// rgzerosum (gauss1.f90:1219-1417) on the device (column-major arrays).
template <class T> void rgzerosum_dev(mmh_ctx* c, T* rhs, const T* def, T dx, int n, double* sums) {
  const i64 n3 = cube(n);
  T* det = wsT<T>(c, "det", n3);
  k_det<T><<<nblk(n3), 256, 0, c->stream>>>(det, def, n, dx);
  LAUNCHCHK();
  reduce<T>(c, det, n3, 0, sums + 0);
  reduce<T>(c, rhs, n3, 0, sums + 1);
  k_rgz_update<T><<<nblk(n3), 256, 0, c->stream>>>(rhs, det, sums, n3);
  LAUNCHCHK();
}

template <class T> void fetch_stats(mmh_ctx* c, Hier<T>& H, mmh_mg_stats* st) {
  if (!st) return;
  std::vector<double> sc(3 * MMH_MAXLEV);
  std::vector<int> g(MMH_MAXLEV + 1);
  HIPCHK(hipMemcpyAsync(sc.data(), H.scal, sizeof(double) * 3 * H.nlev, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(g.data(), H.guard, sizeof(int) * (MMH_MAXLEV + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memset(st, 0, sizeof(*st));
  st->nlev = H.nlev;
  for (int l = 0; l < H.nlev; ++l) {
    st->n[l] = H.n[l];
    st->r1[l] = sc[3 * l];
    st->relaxerr[l] = sc[3 * l + 1];
    st->guard[l] = g[l];
  }
  st->entry_zerosum = g[MMH_MAXLEV];
}

// multigrid entry (multigrid.f90:16-90).  The projection and the entry check act on the caller's
// column-major arrays; the cycle itself runs on sub-lattice copies of u and rhs.
template <class T>
void multigrid_dev(mmh_ctx* c, T* u, T* rhs, const T* defpot, const T* ubig, T dx, int n, int nu, int iopt, int ng,
                   mmh_mg_stats* st) {
  require(is_pow2(n) && n <= 1024, "multigrid: n must be a power of two in [4, 1024]");
  if (ng <= 0) ng = n;
  require(ng >= n && is_pow2(ng), "multigrid: ng must be a power of two >= n");
  require(nu >= 1, "multigrid: nu >= 1");
  const int lsquare = (iopt % 8) == 1;
  const i64 n3 = cube(n);
  T* us = wsT<T>(c, "mg_us", n3);
  T* rs = wsT<T>(c, "mg_rs", n3);
  Hier<T> H;
  build_hier(c, H, "mg", us, rs, defpot, ubig, nu, dx, n, lsquare);
  double* sums = H.scal + 3 * MMH_MAXLEV;
  rgzerosum_dev(c, rhs, defpot, dx, n, sums);
  reduce<T>(c, u, n3, 0, sums + 2);
  k_entry_zerosum<T><<<nblk(n3), 256, 0, c->stream>>>(u, sums + 2, n3, H.guard + MMH_MAXLEV);
  LAUNCHCHK();
  relayout(c, us, u, n, 1, 1);
  relayout(c, rs, rhs, n, 1, 1);
  Sched S;
  S.kind = 0;
  S.ng = ng;
  vcycle(c, H, 0, ng, S);
  relayout(c, u, us, n, 1, 0);
  fetch_stats(c, H, st);
  HIPCHK(hipStreamSynchronize(c->stream));
}

template <class T>
void relax_api(mmh_ctx* c, T* arr, T* resid, const T* rhs, const T* def, const T* u, int nu, T* rinf, T dx, int n, int nrelax,
               int iopt) {
  require(is_pow2(n) && n <= 1024, "relax: n must be a power of two in [4, 1024]");
  require(iopt >= 0 && iopt <= 256, "relax: iopt out of range");
  require(iopt % 8 <= 3, "relax: mod(iopt,8) must be 0..3");
  const i64 n3 = cube(n);
  const int ires = iopt % 8;
  const int lsquare = ((iopt / 8) % 8) == 1;
  T* G = wsT<T>(c, "relaxG", 6 * n3);
  k_metric<T, 1><<<nblk(n3), 256, 0, c->stream>>>(G, def, u, nu, n, dx, lsquare);
  LAUNCHCHK();
  T* as = wsT<T>(c, "relax_as", n3);
  T* bs = wsT<T>(c, "relax_bs", n3);
  T* cs = wsT<T>(c, "relax_cs", cube(n / 2));
  relayout(c, as, arr, n, 1, 1);
  relayout(c, bs, rhs, n, 1, 1);
  double* d = (double*)ws_get(c, "relaxrinf", sizeof(double));
  T* rs = wsT<T>(c, "relaxr", n3);
  const bool want_res = ires == 1 || ires == 3;
  relax_dev(c, as, bs, G, rs, n, dx, nrelax, ires != 0, want_res ? cs : (T*)nullptr, d);
  relayout(c, arr, as, n, 1, 0);
  if (want_res) relayout(c, resid, cs, n / 2, 1, 0);
  double h = 0;
  if (ires == 2 || ires == 3) HIPCHK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rinf) *rinf = (T)h;
}

// ------------------------------------------------------------------------------------------------
// FFT driver
template <class T> const C2<T>* twiddle_table(mmh_ctx* c, int L, int s) {
  const std::string key = "tw" + std::to_string(L) + (s > 0 ? "p" : "m");
  C2<T>* d = wsT<C2<T>>(c, key, std::max(1, L / 2));
  const std::string ukey = key + ":" + std::to_string(sizeof(T));
  if (!c->uploaded.count(ukey)) {
    c->uploaded.insert(ukey);
    std::vector<C2<T>> h(std::max(1, L / 2));
    for (int q = 0; q < L / 2; ++q) {
      double ang = 2.0 * M_PI * (double)q / (double)L;
      h[q].r = (T)cos(ang);
      h[q].i = (T)((double)s * sin(ang));
    }
    HIPCHK(hipMemcpyAsync(d, h.data(), sizeof(C2<T>) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return d;
}

template <class T> void fft3_dev(mmh_ctx* c, C2<T>* z, int d0, int d1, int d2, int s) {
  const int dims[3] = {d0, d1, d2};
  for (int ax = 0; ax < 3; ++ax) {
    const int L = dims[ax];
    if (L == 1) continue;
    const C2<T>* W = twiddle_table<T>(c, L, s);
    const i64 nl = (i64)d0 * d1 * d2 / L;
    const int thr = std::max(64, std::min(256, L / 2));
    k_fft_lines<T><<<(unsigned)nl, thr, 2 * L * sizeof(C2<T>), c->stream>>>(z, W, L, d0, d1, ax);
    LAUNCHCHK();
  }
}

template <class T> __global__ void k_get_speq(C2<T>* sq, const C2<T>* z, int h1, i64 m) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < m) sq[q] = z[(i64)h1 * q];
}

template <class T> void rlft3_dev(mmh_ctx* c, T* data, T* speq, int n, int isign) {
  C2<T>* z = reinterpret_cast<C2<T>*>(data);
  C2<T>* sq = reinterpret_cast<C2<T>*>(speq);
  const int h1 = n / 2;
  const i64 m = (i64)n * n;
  if (isign == 1) {
    fft3_dev<T>(c, z, h1, n, n, 1);
    k_get_speq<T><<<nblk(m), 256, 0, c->stream>>>(sq, z, h1, m);
    LAUNCHCHK();
  }
  C2<T>* sold = wsT<C2<T>>(c, "rlft_sold", m);
  HIPCHK(hipMemcpyAsync(sold, sq, sizeof(C2<T>) * m, hipMemcpyDeviceToDevice, c->stream));
  const T c2i = T(-0.5) * T(isign);
  k_rlft_k0<T><<<nblk(m), 256, 0, c->stream>>>(z, sq, sold, h1, n, n, c2i);
  LAUNCHCHK();
  const int npair = (h1 - 1) / 2;
  if (npair > 0) {
    std::vector<C2<T>> w(h1);
    for (int i1 = 0; i1 < h1; ++i1) {
      double ang = 2.0 * M_PI * (double)i1 / (double)n;
      w[i1].r = (T)cos(ang);
      w[i1].i = (T)((double)isign * sin(ang));
    }
    C2<T>* wd = wsT<C2<T>>(c, std::string("rlft_w") + (isign > 0 ? "p" : "m") + std::to_string(n), h1);
    HIPCHK(hipMemcpyAsync(wd, w.data(), sizeof(C2<T>) * h1, hipMemcpyHostToDevice, c->stream));
    k_rlft_pairs<T><<<nblk((i64)npair * m), 256, 0, c->stream>>>(z, wd, h1, n, n, c2i);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (isign == -1) fft3_dev<T>(c, z, h1, n, n, -1);
}

// Spectral multiplier pass on (data, speq); mode as in k_spectral.
template <class T> void spectral(mmh_ctx* c, T* data, T* speq, int n, int mode) {
  const T pi2 = T(3.14159265358979323846264338328);      // limiter.f90:59
  const T piG = (T)(atan(1.0) * 4.0);                   // gradient.f90:10, reconstruct.f90:9
  T* cs = wsT<T>(c, "spec_cos" + std::to_string(n), 3 * (i64)n);
  if (mode == 0) {
    std::vector<T> h(3 * (size_t)n);
    for (int ii = 0; ii <= n / 2; ++ii) h[ii] = (T)std::cos((T)(T(2) * pi2 * T(ii) / T(n)));
    for (int j = 0; j < n; ++j) h[n + j] = (T)std::cos((T)(T(2) * pi2 * (T(j + 1) - T(1.)) / T(n)));
    for (int k = 0; k < n; ++k) h[2 * n + k] = h[n + k];
    HIPCHK(hipMemcpyAsync(cs, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  const i64 cnt = (i64)(n / 2 + 1) * n * n;
  k_spectral<T><<<nblk(cnt), 256, 0, c->stream>>>(data, speq, n, mode, cs, cs + n, cs + 2 * n, piG);
  LAUNCHCHK();
}

template <class T> void nsmooth_dev(mmh_ctx* c, T* a, int n, int times) {
  const i64 n3 = cube(n);
  T* t = wsT<T>(c, "smooth_tmp", n3);
  for (int r = 0; r < times; ++r) {
    k_smooth27<T><<<nblk(n3), 256, 0, c->stream>>>(t, a, n);
    HIPCHK(hipMemcpyAsync(a, t, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
  }
  LAUNCHCHK();
}

template <class T> __global__ void k_pad(T* __restrict__ dst, const T* __restrict__ src, int n, i64 n3) {
  // (n x n x n) -> leading dimension n+2 (defplim / tmp2 layout, limiter.f90:43)
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const i64 i = p % n, r = p / n;
  dst[i + (i64)(n + 2) * r] = src[p];
}
template <class T> __global__ void k_pad_speq(T* __restrict__ dst, const T* __restrict__ sq, int n) {
  const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (i64)n * n) return;
  dst[(i64)n + (i64)(n + 2) * r] = sq[2 * r];
  dst[(i64)n + 1 + (i64)(n + 2) * r] = sq[2 * r + 1];
}

// Steps 1-3 of calcdefp: the right-hand side handed to multigrid (limiter.f90:134-392).  hb
// receives the per-block [nexpand, detmax, detmin, fluxmax, fluxmin] (valid after a sync).
template <class T>
void limiter_rhs_dev(mmh_ctx* c, T* tmp, const T* def, const T* u, T dtold, T dtaumesh, int nf, int n, T compressmax,
                     std::vector<double>& hb, T* gate = nullptr) {
  const T pi = T(3.14159265358979323846264338328);
  const i64 n3 = cube(n);
  k_mismatch<T><<<nblk(n3), 256, 0, c->stream>>>(tmp, u, nf, n3, dtold);
  LAUNCHCHK();
  nsmooth_dev(c, tmp, n, 3);
  T* B = wsT<T>(c, "limB", 7 * n3);
  k_binv<T><<<nblk(n3), 256, 0, c->stream>>>(B, def, n);
  T* ef = wsT<T>(c, "limE", n3);
  const unsigned nb = nblk(n3);
  double* bst = (double*)ws_get(c, "limstat", sizeof(double) * 5 * nb);
  k_eflux<T><<<nb, 256, 0, c->stream>>>(ef, tmp, B, def, u, nf, n, dtold, dtaumesh, compressmax, pi, bst, gate);
  LAUNCHCHK();
  hb.assign(5 * (size_t)nb, 0.0);
  HIPCHK(hipMemcpyAsync(hb.data(), bst, sizeof(double) * hb.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(tmp, ef, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
  nsmooth_dev(c, tmp, n, 1);
}

void solve_dev(mmh_ctx* c, double* u, double* rhs, const double* defpot, const double* ubig, int n, int nu,
               const mmh_solve_opts* o, mmh_solve_stats* st);

// calcdefp's mesh solve to a tolerance instead of the reference's one V-cycle (limiter.f90:408):
// option "mesh_solve" 1 repeats V(2,2) cycles of the reference V-cycle (divergence guard included,
// 32 + 32 sweeps on 4^3), 2 runs the Krylov solve; both stop at max|rhs - L u| / max|rhs| <=
// 10^-mesh_tol.  FP64 only.
inline void mesh_solve_dev(mmh_ctx* c, double* defp, double* rhs, const double* def, const double* u, int n, int nf,
                           int ng) {
  mmh_solve_opts o;
  std::memset(&o, 0, sizeof(o));
  o.tol = std::pow(10.0, -c->mesh_tol);
  o.max_cycles = c->mesh_solve == 1 ? 100 : 256;
  o.schedule = 1;
  o.nu_pre = 2; o.nu_post = 2;
  o.nu_coarse = c->mesh_solve == 1 ? 32 : 4;
  o.ng = ng;
  o.krylov = c->mesh_solve == 2;
  mmh_solve_stats st;
  solve_dev(c, defp, rhs, def, u, n, nf, &o, &st);
  c->mesh_cycles = st.cycles;
}
inline void mesh_solve_dev(mmh_ctx*, float*, float*, const float*, const float*, int, int, int) {
  throw MmhError(MMH_EINVAL, "calcdefp: mesh_solve (a solve to tolerance) is FP64 only");
}

// calcdefp (limiter.f90:17-673) on the device.
template <class T>
void calcdefp_dev(mmh_ctx* c, T* defp, T* tmp, T* tmp2, const T* def, const T* u, T dtold, T dtaumesh, int nf, int n, T compressmax,
                  int ng, mmh_lim_stats* st) {
  require(is_pow2(n) && n >= 16 && n <= 1024, "calcdefp: n must be a power of two in [16, 1024]");
  require(nf >= 4, "calcdefp: nfluid must be >= 4 (u(2:4) are read)");
  require(compressmax >= T(0), "calcdefp: compressmax < 0 selects the reference's diagnostic-only branch");
  if (dtold < T(1.e-8)) dtold = 1;
  if (ng <= 0) ng = n;
  const T pi = T(3.14159265358979323846264338328);
  const i64 n3 = cube(n);
  std::vector<double> hb;
  limiter_rhs_dev<T>(c, tmp, def, u, dtold, dtaumesh, nf, n, compressmax, hb);
  fill<T>(c, defp, T(0), n3);
  mmh_mg_stats* mst = st ? &st->mg : nullptr;
  if (c->mesh_solve) mesh_solve_dev(c, defp, tmp, def, u, n, nf, ng);
  else multigrid_dev<T>(c, defp, tmp, def, u, T(1.), n, nf, 2, ng, mst);
  nsmooth_dev(c, defp, n, 2);
  const unsigned nb = nblk(n3);
  T* lim = wsT<T>(c, "limL", n3);
  LAUNCHCHK();
  double* bst2 = (double*)ws_get(c, "limstat2", sizeof(double) * 2 * nb);
  k_complim<T><<<nb, 256, 0, c->stream>>>(lim, defp, def, n, dtold, compressmax, pi, bst2);
  LAUNCHCHK();
  std::vector<double> hb2(2 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(hb2.data(), bst2, sizeof(double) * hb2.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  long ncompress = 0;
  double cmax = 0;
  for (unsigned b = 0; b < nb; ++b) { ncompress += (long)hb2[2 * b]; cmax = std::max(cmax, hb2[2 * b + 1]); }
  k_pad<T><<<nblk(n3), 256, 0, c->stream>>>(tmp2, lim, n, n3);
  LAUNCHCHK();
  if (ncompress > 0) {
    T* sq = wsT<T>(c, "limS", 2 * (i64)n * n);
    fill<T>(c, sq, T(0), 2 * (i64)n * n);
    rlft3_dev<T>(c, lim, sq, n, 1);
    spectral<T>(c, lim, sq, n, 0);
    fill<T>(c, lim, T(0), 1);
    rlft3_dev<T>(c, lim, sq, n, -1);
    k_axpy_scale<T><<<nblk(n3), 256, 0, c->stream>>>(defp, lim, n, n3);
    k_scale2<T><<<nblk(n3), 256, 0, c->stream>>>(lim, lim, n, n3);
    k_pad<T><<<nblk(n3), 256, 0, c->stream>>>(tmp2, lim, n, n3);
    k_pad_speq<T><<<nblk((i64)n * n), 256, 0, c->stream>>>(tmp2, sq, n);
    LAUNCHCHK();
  }
  // exit convention of tmp (limiter.f90:621-628,669-671): all cells "hot" -> 1; diagnostics 0
  fill<T>(c, tmp, T(1), n3);
  fill<T>(c, tmp, T(0), 3);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (st) {
    double detmax = 0, detmin = 1, fmaxv = 0, fminv = 0;
    long nexp = 0;
    for (unsigned b = 0; b < nb; ++b) {
      nexp += (long)hb[5 * b];
      detmax = std::max(detmax, hb[5 * b + 1]);
      detmin = std::min(detmin, hb[5 * b + 2]);
      fmaxv = std::max(fmaxv, hb[5 * b + 3]);
      fminv = std::min(fminv, hb[5 * b + 4]);
    }
    st->cmax = cmax; st->detmin = detmin; st->detmax = detmax; st->fluxmin = fminv; st->fluxmax = fmaxv;
    st->nexpand = (int)nexp; st->ncompress = (int)ncompress;
  }
}

template <class T> void relaxing_dev(mmh_ctx* c, T* u, const T* def, const T* defp, T dt, int nf, int n) {
  require(is_pow2(n) && n >= 16 && n <= 1024, "relaxing: n must be a power of two in [16, 1024]");
  require(nf == 5, "relaxing: only nfluid = 5 is supported (relaxing.f90:307-310)");
  const i64 n3 = cube(n);
  T* B = wsT<T>(c, "tvdB", 7 * n3);
  T* cc = wsT<T>(c, "tvdC", n3);
  T* V = wsT<T>(c, "tvdV", 15 * n3);
  T* u1 = wsT<T>(c, "tvdU1", 5 * n3);
  const unsigned nb = nblk(n3);
  k_binv<T><<<nb, 256, 0, c->stream>>>(B, def, n);
  k_cfl<T><<<nb, 256, 0, c->stream>>>(cc, B, defp, u, n);
  k_tvd_flux<T><<<nb, 256, 0, c->stream>>>(V, u, defp, B, cc, n);
  k_tvd_step<T><<<nb, 256, 0, c->stream>>>(u1, u, u, V, cc, n, dt / T(2));
  k_tvd_flux<T><<<nb, 256, 0, c->stream>>>(V, u1, defp, B, cc, n);
  k_tvd_step<T><<<nb, 256, 0, c->stream>>>(u, u, u1, V, cc, n, dt);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(c->stream));
}

// The reconstruction iteration of mmh/main.f90:89-110 with every field resident on the device.
template <class T>
void reconstruct_dev(mmh_ctx* c, T* u, T* def, T* defp, int n, int nf, int niter, T dtold, T dtaumesh, T dt,
                     T compressmax, mmh_iter_stats* st) {
  require(is_pow2(n) && n >= 16 && n <= 1024, "reconstruct: n must be a power of two in [16, 1024]");
  require(nf == 5, "reconstruct: nfluid must be 5 (relaxing.f90:307-310)");
  require(niter >= 0, "reconstruct: niter < 0");
  const i64 n3 = cube(n);
  T* tmp = wsT<T>(c, "rc_tmp", n3);
  T* tmp2 = wsT<T>(c, "rc_tmp2", (i64)(n + 2) * n * n);
  const unsigned nb = nblk(n3, RB * RPT);
  double* mom = (double*)ws_get(c, "rc_mom", sizeof(double) * 6 * nb);
  std::vector<double> hm(6 * (size_t)nb);
  for (int it = 0; it < niter; ++it) {
    mmh_iter_stats* s = st ? st + it : nullptr;
    calcdefp_dev<T>(c, defp, tmp, tmp2, def, u, dtold, dtaumesh, nf, n, compressmax, n, s ? &s->lim : nullptr);
    k_add<T><<<nblk(n3), 256, 0, c->stream>>>(def, defp, n3);
    LAUNCHCHK();
    relaxing_dev<T>(c, u, def, defp, dt, nf, n);
    if (!s) continue;
    k_iter_moments<T><<<nb, RB, 0, c->stream>>>(mom, defp, def, u, nf, n3);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(hm.data(), mom, sizeof(double) * hm.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double a0 = 0, a1 = 0, a2 = 0, mn = 1e300, mx = -1e300, ng = 0;
    for (unsigned b = 0; b < nb; ++b) {
      a0 += hm[6 * b]; a1 += hm[6 * b + 1]; a2 += hm[6 * b + 2];
      mn = std::min(mn, hm[6 * b + 3]); mx = std::max(mx, hm[6 * b + 4]); ng += hm[6 * b + 5];
    }
    s->iter = it + 1;
    s->defp_sigma = std::sqrt(a0 / (double)n3);
    s->def_sigma = std::sqrt(a1 / (double)n3);
    s->u_sigma = std::sqrt(a2 / (double)n3);
    s->u_min = mn; s->u_max = mx; s->u_negative = (long long)ng;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
}

template <class T> void spectral_out(mmh_ctx* c, const T* phi, T* out, int n, int mode) {
  const i64 n3 = cube(n);
  T* d = wsT<T>(c, "specD", n3);
  T* sq = wsT<T>(c, "specS", 2 * (i64)n * n);
  HIPCHK(hipMemcpyAsync(d, phi, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
  fill<T>(c, sq, T(0), 2 * (i64)n * n);
  rlft3_dev<T>(c, d, sq, n, 1);
  spectral<T>(c, d, sq, n, mode);
  rlft3_dev<T>(c, d, sq, n, -1);
  k_scale2<T><<<nblk(n3), 256, 0, c->stream>>>(out, d, n, n3);
  LAUNCHCHK();
}

// ------------------------------------------------------------------------------------------------
// Lognormal generator kernels
__device__ __forceinline__ uint64_t smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__global__ void k_white(C2<double>* z, i64 n3, uint64_t key) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n3) return;
  uint64_t h1 = smix(key + 2 * (uint64_t)q), h2 = smix(key + 2 * (uint64_t)q + 1);
  double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;
  double u2 = (double)(h2 >> 11) * 0x1.0p-53;
  z[q].r = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  z[q].i = 0.0;
}
__global__ void k_pk(C2<double>* z, int n, double slope, double ks) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cube(n)) return;
  const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((i64)n * n));
  double mi = i <= n / 2 ? i : i - n, mj = j <= n / 2 ? j : j - n, mk = k <= n / 2 ? k : k - n;
  double kx = 2.0 * M_PI * mi / n, ky = 2.0 * M_PI * mj / n, kz = 2.0 * M_PI * mk / n;
  double kk = sqrt(kx * kx + ky * ky + kz * kz);
  double a = kk > 0 ? sqrt(pow(kk, slope) * exp(-(kk / ks) * (kk / ks))) : 0.0;
  z[p].r = z[p].r * a;
  z[p].i = z[p].i * a;
}
__global__ void k_re_sq(double* g, double* g2, const C2<double>* z, i64 n3) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n3) return;
  g[q] = z[q].r;
  g2[q] = z[q].r * z[q].r;
}
__global__ void k_expmap(double* g, const double* sums, i64 n3, double sigma) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n3) return;
  const double var = sums[0] / (double)n3;
  const double sc = sigma / sqrt(var);
  g[q] = exp(g[q] * sc - 0.5 * sigma * sigma);
}
__global__ void k_divmean(double* rho, const double* g, const double* sums, i64 n3) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n3) return;
  rho[q] = g[q] / (sums[1] / (double)n3);
}

uint64_t smix_h(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void lognormal_dev(mmh_ctx* c, double* rho, int n, double sigma, double slope, double ks_frac, uint64_t seed) {
  require(is_pow2(n) && n <= 2048, "lognormal: n must be a power of two");
  const i64 n3 = cube(n);
  C2<double>* z = wsT<C2<double>>(c, "gen_z", n3);
  double* g2 = wsT<double>(c, "gen_g2", n3);
  double* sums = (double*)ws_get(c, "gen_sums", 2 * sizeof(double));
  k_white<<<nblk(n3), 256, 0, c->stream>>>(z, n3, smix_h(seed ^ 0x6A09E667F3BCC909ULL));
  LAUNCHCHK();
  fft3_dev<double>(c, z, n, n, n, -1);
  k_pk<<<nblk(n3), 256, 0, c->stream>>>(z, n, slope, ks_frac * M_PI);
  LAUNCHCHK();
  fft3_dev<double>(c, z, n, n, n, 1);
  k_re_sq<<<nblk(n3), 256, 0, c->stream>>>(rho, g2, z, n3);
  LAUNCHCHK();
  reduce<double>(c, g2, n3, 0, sums);
  k_expmap<<<nblk(n3), 256, 0, c->stream>>>(rho, sums, n3, sigma);
  LAUNCHCHK();
  reduce<double>(c, rho, n3, 0, sums + 1);
  k_divmean<<<nblk(n3), 256, 0, c->stream>>>(rho, rho, sums, n3);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(c->stream));
}

// ------------------------------------------------------------------------------------------------
// Solve to tolerance.  FP64: repeated V-cycles, the relative residual read after each.  Mixed:
// r = rhs - L u in FP64; an FP32 V-cycle solves L e = r from e = 0 (its 4^3 level in FP64); u += e.
template <class S, class D> __global__ void k_convert(D* __restrict__ dst, const S* __restrict__ src, i64 n) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = (D)src[q];
}
// The same over a whole coefficient array (all its fields) in one launch: a grid-stride loop over a
// bounded grid, so any n fits.
template <class S, class D> __global__ void k_convert_all(D* __restrict__ dst, const S* __restrict__ src, i64 n) {
  for (i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (i64)gridDim.x * blockDim.x) dst[q] = (D)src[q];
}
template <class S, class D> void convert_all(mmh_ctx* c, D* dst, const S* src, i64 n) {
  const i64 nb = std::min<i64>((n + 255) / 256, 16384);  // bounded grid; the loop covers the rest
  k_convert_all<S, D><<<(unsigned)std::max<i64>(nb, 1), 256, 0, c->stream>>>(dst, src, n);
  LAUNCHCHK();
}
__global__ void k_add_f2d(double* __restrict__ u, const float* __restrict__ e, i64 n) {
  const i64 q = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) u[q] = u[q] + (double)e[q];
}

template <class T> double read_scalar(mmh_ctx* c, const double* d) {
  double h = 0;
  HIPCHK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return h;
}

// FP32 V-cycle whose 4^3 level runs in FP64 (the coarsest solve stays FP64).
void vcycle_mixed(mmh_ctx* c, Hier<float>& Hf, Hier<double>& Hd, int l, int ng_top, const Sched& S) {
  const int n = Hf.n[l];
  if (coarse_cycle_applies(c, Hf, l, S)) {
    CoarseCycle<float> C = coarse_cycle_of(Hf, l, S, ng_top);
    const int lb = Hd.nlev - 1;
    C.mixed = 1;
    C.du = Hd.u[lb];
    C.drhs = Hd.rhs[lb];
    C.dG = Hd.G[lb];
    C.dscal = Hd.scal + 3 * lb;
    C.dguard = Hd.guard + lb;
    C.ddx = Hd.dx[lb];
    C.dA = Hd.A[lb];
    if (c->coarse_lds) {
      stage_levels(C);
      k_coarse_cycle_lds<float><<<1, CW, 0, c->stream>>>(C);
    } else {
      k_coarse_cycle<float><<<1, CW, 0, c->stream>>>(C);
    }
    LAUNCHCHK();
    return;
  }
  if (n == 4) {
    const i64 n3 = 64;
    k_convert<float, double><<<1, 64, 0, c->stream>>>(Hd.rhs[l], Hf.rhs[l], n3);
    fill<double>(c, Hd.u[l], 0.0, n3);
    LAUNCHCHK();
    vcycle<double>(c, Hd, l, ng_top, S);
    k_convert<double, float><<<1, 64, 0, c->stream>>>(Hf.u[l], Hd.u[l], n3);
    LAUNCHCHK();
    return;
  }
  const i64 n3 = cube(n);
  double* sc = Hf.scal + 3 * l;
  const bool zero = c->zero_start && (l > 0 || S.top_zero);
  if (Hf.Ab[l]) gal_relax<float, bf16x2>(c, Hf.u[l], Hf.rhs[l], Hf.Ab[l], Hf.r[l], n, S.pre(n), Hf.rhs[l + 1], zero);
  else if (Hf.A[l]) gal_relax<float>(c, Hf.u[l], Hf.rhs[l], Hf.A[l], Hf.r[l], n, S.pre(n), Hf.rhs[l + 1], zero);
  else if (Hf.Gb[l]) {
    c->slab_now = (l == 0 && S.top_zero) ? c->sweep_slabs : 0;
    relax_dev<float, bf16>(c, Hf.u[l], Hf.rhs[l], Hf.Gb[l], Hf.r[l], n, Hf.dx[l], S.pre(n), S.guard, Hf.rhs[l + 1],
                           sc + 0, false, S.delta, zero && !c->slab_now);
    c->slab_now = 0;
  }
  else relax_dev(c, Hf.u[l], Hf.rhs[l], Hf.G[l], Hf.r[l], n, Hf.dx[l], S.pre(n), S.guard, Hf.rhs[l + 1], sc + 0, false,
                 S.delta, zero);
  fill<float>(c, Hf.u[l + 1], 0.f, cube(n / 2));
  vcycle_mixed(c, Hf, Hd, l + 1, ng_top, S);
  k_prolong<float, 1><<<nblk(cube(n / 2)), 256, 0, c->stream>>>(Hf.u[l], Hf.u[l + 1], n, 1);
  LAUNCHCHK();
  if (Hf.Ab[l]) gal_sweeps<float, bf16x2>(c, Hf.u[l], Hf.rhs[l], Hf.Ab[l], n, S.post(n), S.symmetric);
  else if (Hf.A[l]) gal_sweeps<float>(c, Hf.u[l], Hf.rhs[l], Hf.A[l], n, S.post(n), S.symmetric);
  else if (Hf.Gb[l]) {
    c->slab_now = (l == 0 && S.top_zero) ? c->sweep_slabs : 0;
    relax_dev<float, bf16>(c, Hf.u[l], Hf.rhs[l], Hf.Gb[l], Hf.r[l], n, Hf.dx[l], S.post(n), S.guard, (float*)nullptr,
                           sc + 1, S.symmetric, S.delta);
    c->slab_now = 0;
  }
  else relax_dev(c, Hf.u[l], Hf.rhs[l], Hf.G[l], Hf.r[l], n, Hf.dx[l], S.post(n), S.guard, (float*)nullptr, sc + 1,
                 S.symmetric, S.delta);
  if (S.defer_top_mean && l == 0) {
    if (!S.sum_by_dots) reduce<float>(c, Hf.u[l], n3, 0, sc + 2);
    return;
  }
  reduce<float>(c, Hf.u[l], n3, 0, sc + 2);
  k_zerosum_guard<float><<<nblk(n3), 256, 0, c->stream>>>(Hf.u[l], sc + 2, sc, n3, (n == ng_top || !S.guard) ? 0 : 1,
                                                          Hf.guard + l);
  LAUNCHCHK();
}

// FP64 -> bf16 copy of n values (the coefficients of the 16-bit preconditioner).
__global__ void k_to_bf16(bf16* __restrict__ dst, const double* __restrict__ src, i64 n) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[p] = to_bf16(src[p]);
}
// The 27 coefficient fields of m3 points (FP64) into 14 words per point: word f of point p holds
// fields 2f (low half) and 2f + 1 (high half; 0 for f = 13), each rounded as k_to_bf16 rounds.
__global__ void k_to_bf16x2(bf16x2* __restrict__ dst, const double* __restrict__ src, i64 m3, int f) {
  const i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m3) return;
  const unsigned lo = to_bf16(src[(i64)(2 * f) * m3 + p]).b;
  const unsigned hi = 2 * f + 1 < 27 ? to_bf16(src[(i64)(2 * f + 1) * m3 + p]).b : 0u;
  dst[(i64)f * m3 + p].w = lo | (hi << 16);
}
void convert_bf16x2(mmh_ctx* c, bf16x2* dst, const double* src, i64 m3) {
  require(m3 <= ((i64)1 << 30), "convert_bf16x2: level too large for one launch per word");
  for (int f = 0; f < 14; ++f) k_to_bf16x2<<<nblk(m3), 256, 0, c->stream>>>(dst, src, m3, f);
  LAUNCHCHK();
}
void convert_bf16(mmh_ctx* c, bf16* dst, const double* src, i64 n) {
  constexpr i64 CH = (i64)1 << 30;  // one launch per 2^30 values (32-bit dispatch sizes)
  for (i64 o = 0; o < n; o += CH) {
    const i64 m = std::min(CH, n - o);
    k_to_bf16<<<nblk(m), 256, 0, c->stream>>>(dst + o, src + o, m);
  }
  LAUNCHCHK();
}

// The FP32 preconditioner's hierarchy.  bf16 (option "precond_bf16", levels of n >= 32): the
// metric of the finest level and the Galerkin operators are stored in 16 bits, read and widened
// by the FP32 sweeps and residuals; smaller levels keep FP32 copies (the fused coarse cycle reads
// those).
void build_fp32(mmh_ctx* c, Hier<float>& F, Hier<double>& H, bool bf = false) {
  F.nlev = H.nlev;
  F.scal = (double*)ws_get(c, "sfscal", sizeof(double) * 4 * MMH_MAXLEV);
  F.guard = (int*)ws_get(c, "sfguard", sizeof(int) * (MMH_MAXLEV + 1));
  for (int l = 0; l < H.nlev; ++l) {
    const int m = H.n[l];
    const std::string t = "sf" + std::to_string(l);
    F.n[l] = m; F.dx[l] = (float)H.dx[l]; F.nu[l] = 1;
    F.u[l] = wsT<float>(c, t + "u", cube(m));
    F.rhs[l] = wsT<float>(c, t + "rhs", cube(m));
    F.r[l] = wsT<float>(c, t + "r", cube(m));
    F.def[l] = nullptr; F.u1[l] = nullptr;
    F.A[l] = nullptr; F.Ab[l] = nullptr; F.Gb[l] = nullptr; F.G[l] = nullptr;
    const bool b16 = bf && m >= 32;
    if (H.A[l]) {
      if (b16) {
        bf16x2* a = (bf16x2*)ws_get(c, t + "Ab2", sizeof(bf16x2) * 14 * cube(m));
        convert_bf16x2(c, a, H.A[l], cube(m));
        F.Ab[l] = a;
      } else {
        F.A[l] = wsT<float>(c, t + "A", 27 * cube(m));
        convert_all<double, float>(c, F.A[l], H.A[l], 27 * cube(m));
      }
    }
    if (b16 && !H.A[l]) {
      bf16* g = (bf16*)ws_get(c, t + "Gb", sizeof(bf16) * 6 * cube(m));
      convert_bf16(c, g, H.G[l], 6 * cube(m));
      F.Gb[l] = g;
    } else if (!b16) {
      F.G[l] = wsT<float>(c, t + "G", 6 * cube(m));
      convert_all<double, float>(c, F.G[l], H.G[l], 6 * cube(m));
    }
    LAUNCHCHK();
  }
}

double dot_dev(mmh_ctx* c, const double* a, const double* b, i64 n, double* dst) {
  const unsigned nb = nblk(n, RB);
  double* bs = (double*)ws_get(c, "dot_bs", sizeof(double) * nb);
  k_dot<double><<<nb, RB, 0, c->stream>>>(bs, a, b, n);
  LAUNCHCHK();
  reduce<double>(c, bs, nb, 0, dst);
  return read_scalar<double>(c, dst);
}

// Flexible preconditioned conjugate gradients on L x = b (x, b sub-lattice FP64).  The
// preconditioner is one symmetric V-cycle from z = 0 (colours reversed in post-smoothing, no
// divergence guard), in FP64 or, with F, in FP32 with its 4^3 level in FP64.  beta is the
// Polak-Ribiere form z.(r - r_old) / z_old.r_old, which tolerates the FP32 preconditioner's rounding.
void pcg(mmh_ctx* c, Hier<double>& H, Hier<float>* F, double* x, double* b, int n, const Sched& S,
         const mmh_solve_opts* o, mmh_solve_stats* st, double rden, double* sums) {
  const i64 n3 = cube(n);
  // With the FP32 preconditioner (option "dots32"), the dot products z.r and z.r_old that set alpha
  // and beta read r and r_old rounded to FP32, as the preconditioner itself receives r: r_old is
  // kept in FP32 only (the r update writes it), 12 B per point less per iteration.  z comes from an
  // FP32 V-cycle, so the products carry rounding of the same order either way; the residual
  // recurrence and the stopping test stay FP64.
  const bool d32 = F && c->dots32;
  double* r = wsT<double>(c, "cg_r", n3);
  double* ro = d32 ? nullptr : wsT<double>(c, "cg_ro", n3);
  double* p = wsT<double>(c, "cg_p", n3);
  double* q = wsT<double>(c, "cg_q", n3);
  // z = M r: FP64, or with the FP32 preconditioner the FP32 V-cycle's own solution array, read
  // (widened exactly) by the dot products and the direction update; r reaches the FP32 V-cycle as
  // its right-hand side, rounded by the r update itself
  double* z = F ? nullptr : wsT<double>(c, "cg_z", n3);
  float* z32 = F ? F->u[0] : nullptr;
  float* r32 = F ? F->rhs[0] : nullptr;
  float* ro32 = d32 ? wsT<float>(c, "cg_ro32", n3) : nullptr;
  const unsigned nb = nblk(n3, RB);
  double* bm = (double*)ws_get(c, "cg_bm", sizeof(double) * nb);
  double* bs = (double*)ws_get(c, "cg_bs", sizeof(double) * nb);
  double* bs2 = (double*)ws_get(c, "cg_bs2", sizeof(double) * nb);
  // The preconditioner is device work only: several hundred launches per V-cycle, most of them on
  // the 4^3..32^3 levels.  The first call runs eagerly (it sizes every workspace it uses); the
  // second is captured and every later call replays the graph (option "pcg_graph").
  std::unique_ptr<StreamSwap> own_stream;
  if (c->pcg_graph && c->stream == nullptr) own_stream.reset(new StreamSwap(c));
  Replay graph(c);
  int ncalls = 0;
  Hier<double> P = H;
  P.u[0] = z;
  P.rhs[0] = r;
  // The V-cycle's last step, removing the mean of z, is left to z's readers (the dot products and
  // the direction update subtract it as they read, with the expression of k_zerosum_guard), which
  // saves a pass over z; not where the finest level runs inside the one-workgroup coarse cycle.
  Sched Sz = S;
  Sz.top_zero = true;  // z is cleared before each V-cycle
  Sz.defer_top_mean = n > 4 && !(F ? coarse_cycle_applies(c, *F, 0, S) : coarse_cycle_applies(c, H, 0, S));
  const double* zsum = Sz.defer_top_mean ? (F ? F->scal + 2 : H.scal + 2) : nullptr;
  // With dots32 the mixed solve's dot-product pass also forms z's sum (k_dot5): z.r with the mean
  // removed is z.r - (sum z / n) sum r, r and r_old being mean-free up to rounding, and the V-cycle
  // skips its own pass over z for that sum.
  const bool sum5 = d32 && Sz.defer_top_mean;
  Sz.sum_by_dots = sum5;
  const unsigned nd5 = nblk(n3, RB * RPT);
  double* d5 = sum5 ? (double*)ws_get(c, "cg_d5", 5 * sizeof(double)) : nullptr;
  double* bs3 = sum5 ? (double*)ws_get(c, "cg_bs3", sizeof(double) * nd5) : nullptr;
  double* bs4 = sum5 ? (double*)ws_get(c, "cg_bs4", sizeof(double) * nd5) : nullptr;
  double* bs5 = sum5 ? (double*)ws_get(c, "cg_bs5", sizeof(double) * nd5) : nullptr;
  auto precond_work = [&] {
    if (!F) {
      fill<double>(c, z, 0.0, n3);
      vcycle<double>(c, P, 0, n, Sz);
    } else {
      fill<float>(c, z32, 0.f, n3);
      vcycle_mixed(c, *F, H, 0, n, Sz);
    }
  };
  auto precond = [&] {
    if (!c->pcg_graph || ncalls++ == 0) {
      precond_work();
      return;
    }
    if (!graph.exec) graph.capture(precond_work);
    graph.launch();
  };
  // z.r and z.ro of the current z (deterministic block sums, one read-back for both)
  auto dots = [&](const double* rr, const double* rro, double& zr, double& zro) {
    const unsigned nd = nblk(n3, RB * RPT);
    if (sum5) {
      k_dot5<float, float><<<nd, RB, 0, c->stream>>>(bs3, bs, bs2, bs4, bs5, z32, r32, rr == rro ? r32 : ro32, n3);
      LAUNCHCHK();
      reduce<double>(c, bs3, nd, 0, d5 + 0);
      reduce<double>(c, bs, nd, 0, d5 + 1);
      reduce<double>(c, bs2, nd, 0, d5 + 2);
      reduce<double>(c, bs4, nd, 0, d5 + 3);
      reduce<double>(c, bs5, nd, 0, d5 + 4);
      reduce<double>(c, bs3, nd, 0, F->scal + 2);  // z's sum where the direction update reads it
      double h[5];
      HIPCHK(hipMemcpyAsync(h, d5, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      const double m = h[0] / (double)n3;
      zr = h[1] - m * h[3];
      zro = h[2] - m * h[4];
      return;
    }
    if (d32) k_dot2<float, float><<<nd, RB, 0, c->stream>>>(bs, bs2, z32, r32, rr == rro ? r32 : ro32, n3, zsum);
    else if (F) k_dot2<float><<<nd, RB, 0, c->stream>>>(bs, bs2, z32, rr, rro, n3, zsum);
    else k_dot2<double><<<nd, RB, 0, c->stream>>>(bs, bs2, z, rr, rro, n3, zsum);
    LAUNCHCHK();
    reduce<double>(c, bs, nd, 0, sums + 8);
    reduce<double>(c, bs2, nd, 0, sums + 9);
    double h[2];
    HIPCHK(hipMemcpyAsync(h, sums + 8, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    zr = h[0];
    zro = h[1];
  };
  // r = b - L x.  From x = 0 (a cold start) that is b itself (every product with a zero neighbour
  // is zero; only the sign of a zero entry could differ): a max-norm of x, a quarter of the
  // operator's bytes, decides.
  reduce<double>(c, x, n3, 1, sums + 7);
  if (read_scalar<double>(c, sums + 7) == 0.0)
    HIPCHK(hipMemcpyAsync(r, b, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  else
    op_resid<double>(c, r, bm, x, b, H.G[0], n, 1.0);
  if (F) k_convert<double, float><<<nblk(n3), 256, 0, c->stream>>>(r32, r, n3);
  precond();
  double rz, unused;
  dots(r, r, rz, unused);
  // definiteness monitor: p.Lp < 0 and z.r < 0 on a negative definite operator and preconditioner
  int nonneg_pq = 0, nonneg_zr = rz >= 0.0 ? 1 : 0;
  auto indefinite_note = [&] {
    if (st) { st->nonneg_pLp = nonneg_pq; st->nonneg_zr = nonneg_zr; }
    if (!nonneg_pq && !nonneg_zr) return std::string();
    return " (p.Lp >= 0 in " + std::to_string(nonneg_pq) + " and z.r >= 0 in " + std::to_string(nonneg_zr) +
           " iterations: the discrete operator or its preconditioner is not definite)";
  };
  if (F) k_cg_dir<float><<<nblk(n3), 256, 0, c->stream>>>(nullptr, p, z32, 0.0, 0.0, n3, zsum, 0);
  else k_cg_dir<double><<<nblk(n3), 256, 0, c->stream>>>(nullptr, p, z, 0.0, 0.0, n3, zsum, 0);
  LAUNCHCHK();
  for (int it = 0; it < o->max_cycles; ++it) {
    const unsigned nq = op_apply<double>(c, q, bs, p, H.G[0], n, 1.0);
    reduce<double>(c, bs, nq, 0, sums + 5);
    const double pq = read_scalar<double>(c, sums + 5);
    if (pq >= 0.0) ++nonneg_pq;
    const double alpha = rz / pq;
    if (!std::isfinite(alpha)) {
      // p.Lp = 0 (or not a number): the recurrence has broken down; keep the iterate as it is
      if (st) { st->cycles = it; st->nonneg_pLp = nonneg_pq; st->nonneg_zr = nonneg_zr; }
      throw MmhError(MMH_ENOCONV, "solve: breakdown, p.Lp = " + std::to_string(pq));
    }
    k_cg_update<<<nb, RB, 0, c->stream>>>(r, ro, q, alpha, bm, n3, r32, ro32);
    LAUNCHCHK();
    reduce<double>(c, bm, nb, 1, sums + 6);
    const double rr = read_scalar<double>(c, sums + 6) / rden;
    if (st) { st->relres[it] = rr; st->cycles = it + 1; }
    if (rr <= o->tol || it + 1 == o->max_cycles) {
      // the iterate's last update (the others ride on the direction update)
      k_axpy<<<nblk(n3), 256, 0, c->stream>>>(x, p, alpha, n3);
      LAUNCHCHK();
      if (rr <= o->tol) break;
      const std::string note = indefinite_note();
      throw MmhError(MMH_ENOCONV, "solve: tolerance not reached" + note);
    }
    precond();
    double rz_new, zro;
    dots(r, ro, rz_new, zro);
    if (rz_new >= 0.0) ++nonneg_zr;
    const double beta = (rz_new - zro) / rz;
    if (F) k_cg_dir<float><<<nblk(n3), 256, 0, c->stream>>>(x, p, z32, alpha, beta, n3, zsum);
    else k_cg_dir<double><<<nblk(n3), 256, 0, c->stream>>>(x, p, z, alpha, beta, n3, zsum);
    LAUNCHCHK();
    rz = rz_new;
  }
  // report the true residual of the final iterate and remove its mean
  const unsigned nr = op_resid<double>(c, (double*)nullptr, bm, x, b, H.G[0], n, 1.0);
  reduce<double>(c, bm, nr, 1, sums + 6);
  if (st && st->cycles > 0) st->relres[st->cycles - 1] = read_scalar<double>(c, sums + 6) / rden;
  (void)indefinite_note();
  reduce<double>(c, x, n3, 0, sums + 2);
  k_zerosum_guard<double><<<nblk(n3), 256, 0, c->stream>>>(x, sums + 2, H.scal, n3, 0, nullptr);
  LAUNCHCHK();
}

// Mixed precision by iterative refinement: the FP64 residual r = b - L x of the current iterate
// (sub-lattice FP64), an FP32 flexible PCG on L e = r (FP32 operator, vectors and symmetric FP32
// V-cycle preconditioner with the FP64 4^3 level; FP64 dot products) to a relative reduction of
// 1e-4, x += e in FP64; until the FP64 residual meets the tolerance.
template <class T> double dot_t(mmh_ctx* c, const T* a, const T* b, i64 n, double* dst) {
  const unsigned nb = nblk(n, RB);
  double* bs = (double*)ws_get(c, "dot_bs", sizeof(double) * nb);
  k_dot<T><<<nb, RB, 0, c->stream>>>(bs, a, b, n);
  LAUNCHCHK();
  reduce<double>(c, bs, nb, 0, dst);
  return read_scalar<double>(c, dst);
}

void refine_fp32(mmh_ctx* c, Hier<double>& H, Hier<float>& F, double* x, double* b, int n, const Sched& S,
                 const mmh_solve_opts* o, mmh_solve_stats* st, double rden, double* sums) {
  const i64 n3 = cube(n);
  const double inner_tol = 1e-4;
  double* r64 = H.r[0];
  float* e = wsT<float>(c, "rf_e", n3);
  float* r = wsT<float>(c, "rf_r", n3);
  float* ro = wsT<float>(c, "rf_ro", n3);
  float* p = wsT<float>(c, "rf_p", n3);
  float* q = wsT<float>(c, "rf_q", n3);
  float* z = wsT<float>(c, "rf_z", n3);
  const unsigned nb = nblk(n3, RB), ns = stencil_grid(n, 1);
  double* bm = (double*)ws_get(c, "rf_bm", sizeof(double) * nb);
  double* bs = (double*)ws_get(c, "rf_bs", sizeof(double) * nb);
  Hier<float> P = F;
  P.u[0] = z;
  P.rhs[0] = r;
  auto precond = [&] {
    fill<float>(c, z, 0.f, n3);
    vcycle_mixed(c, P, H, 0, n, S);
  };
  int total = 0, outer = 0;
  for (;; ++outer) {
    k_resid<double, 1><<<ns, RB, 0, c->stream>>>(r64, bm, x, b, H.G[0], n, 1.0);
    LAUNCHCHK();
    reduce<double>(c, bm, ns, 1, sums + 6);
    const double rmax = read_scalar<double>(c, sums + 6);
    const double rr = rmax / rden;
    if (st && total > 0) st->relres[total - 1] = rr;
    if (rr <= o->tol) break;
    if (total >= o->max_cycles) throw MmhError(MMH_ENOCONV, "solve: tolerance not reached");
    k_convert<double, float><<<nblk(n3), 256, 0, c->stream>>>(r, r64, n3);
    fill<float>(c, e, 0.f, n3);
    LAUNCHCHK();
    precond();
    double rz = dot_t<float>(c, r, z, n3, sums + 4);
    HIPCHK(hipMemcpyAsync(p, z, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->stream));
    while (total < o->max_cycles) {
      k_apply_dot<float, 1, true><<<ns, RB, 0, c->stream>>>(q, bs, p, F.G[0], n, 1.f);
      LAUNCHCHK();
      reduce<double>(c, bs, ns, 0, sums + 5);
      const double alpha = rz / read_scalar<double>(c, sums + 5);
      k_cg_update_t<float><<<nb, RB, 0, c->stream>>>(e, r, ro, p, q, alpha, bm, n3);
      LAUNCHCHK();
      reduce<double>(c, bm, nb, 1, sums + 7);
      const double rin = read_scalar<double>(c, sums + 7) / rmax;
      ++total;
      if (st) { st->relres[total - 1] = rin * rr; st->cycles = total; }
      if (rin <= inner_tol) break;
      precond();
      const double rz_new = dot_t<float>(c, z, r, n3, sums + 4);
      const double zro = dot_t<float>(c, z, ro, n3, sums + 4);
      const double beta = (rz_new - zro) / rz;
      k_cg_dir_t<float><<<nblk(n3), 256, 0, c->stream>>>(p, z, beta, n3);
      LAUNCHCHK();
      rz = rz_new;
    }
    k_add_f2d<<<nblk(n3), 256, 0, c->stream>>>(x, e, n3);
    LAUNCHCHK();
  }
  reduce<double>(c, x, n3, 0, sums + 2);
  k_zerosum_guard<double><<<nblk(n3), 256, 0, c->stream>>>(x, sums + 2, H.scal, n3, 0, nullptr);
  LAUNCHCHK();
}

void solve_dev(mmh_ctx* c, double* u, double* rhs, const double* defpot, const double* ubig, int n, int nu,
               const mmh_solve_opts* o, mmh_solve_stats* st) {
  require(is_pow2(n) && n <= 1024, "solve: n must be a power of two in [4, 1024]");
  require(o != nullptr, "solve: opts required");
  require(o->max_cycles >= 1 && o->max_cycles <= 256, "solve: max_cycles in [1, 256]");
  Sched S;
  S.kind = o->schedule;
  S.ng = o->ng > 0 ? o->ng : n;
  S.nu_pre = o->nu_pre; S.nu_post = o->nu_post; S.nu_coarse = o->nu_coarse;
  if (S.kind == 1) require(S.nu_pre >= 1 && S.nu_post >= 1 && S.nu_coarse >= 1, "solve: sweep counts >= 1");
  if (o->krylov) {
    require(S.kind == 1 && S.nu_pre == S.nu_post, "solve: the Krylov solve needs a fixed symmetric schedule");
    S.guard = false;
    S.symmetric = true;
    S.ng = n;
    S.nu_top = c->nu_top;
  }
  S.delta = o->mixed != 0 && c->mixed_delta;
  const i64 n3 = cube(n);
  double* us = wsT<double>(c, "sd_us", n3);
  double* rs = wsT<double>(c, "sd_rs", n3);
  Hier<double> H;
  build_hier(c, H, "sd", us, rs, defpot, ubig, nu, 1.0, n, 0);
  // The mixed Krylov solve keeps no FP64 preconditioned residual: an FP64 solve's buffer for it
  // (N^3 words, 8 GiB at 1024^3) goes before the Galerkin operators are sized.
  const bool d32 = o->krylov && o->mixed == 1 && c->dots32;
  if (o->krylov && o->mixed) ws_free(c, "cg_z:8");
  if (d32) ws_free(c, "cg_ro:8");  // its FP32 rounding replaces it
  const bool gal = o->krylov && c->galerkin && !S.delta &&
                   galerkin_fits(c, H, "sd", o->mixed != 0, o->mixed == 1 && c->precond_bf16, d32);
  if (gal) {
    try {
      build_galerkin(c, H, "sd");
    } catch (const MmhError& e) {
      if (e.code != MMH_ENOMEM) throw;
      // out of memory after all: release what was built and use the rebuilt operators
      for (int l = 1; l < H.nlev; ++l) {
        ws_free(c, "sd" + std::to_string(l) + "A:8");
        H.A[l] = nullptr;
      }
    }
  }
  double* sums = H.scal + 3 * MMH_MAXLEV;
  rgzerosum_dev<double>(c, rhs, defpot, 1.0, n, sums);
  reduce<double>(c, rhs, n3, 1, sums + 3);
  relayout(c, us, u, n, 1, 1);
  relayout(c, rs, rhs, n, 1, 1);
  const double rn = read_scalar<double>(c, sums + 3);
  if (st) { std::memset(st, 0, sizeof(*st)); st->rhs_norm = rn; }
  const double rden = rn > 0 ? rn : 1.0;
  auto finish = [&] {
    relayout(c, u, us, n, 1, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
  };
  if (st) st->galerkin = H.A[1] ? 1 : 0;
  if (rn == 0.0) {
    // L u = 0 has the mean-free solution 0
    fill<double>(c, u, 0.0, n3);
    HIPCHK(hipStreamSynchronize(c->stream));
    return;
  }
  if (o->krylov) {
    Hier<float> F;
    if (o->mixed) {
      build_fp32(c, F, H, o->mixed == 1 && c->precond_bf16);
      // The FP32 cycle reads its own copies; only the FP64 4^3 level's operator is read in FP64.
      // The FP64 operators of the other levels stay cached for the next solve, which rounds its
      // copies from them again: freeing them would leave it to rebuild tens of GiB at 1024^3, where
      // the memory may then be short and the solve fall back to the rebuilt operators.  They go
      // only if the Krylov vectors, still to be allocated, would not fit beside them.
      size_t freeb = 0, total = 0;
      HIPCHK(hipMemGetInfo(&freeb, &total));
      size_t vec = 0;
      for (const char* k : {"cg_r:8", "cg_p:8", "cg_q:8"}) vec += ws_need(c, k, (size_t)n3 * sizeof(double));
      vec += d32 ? ws_need(c, "cg_ro32:4", (size_t)n3 * sizeof(float)) : ws_need(c, "cg_ro:8", (size_t)n3 * sizeof(double));
      if (vec > 0 && freeb < vec + ((size_t)1 << 30))
        for (int l = 1; l + 1 < H.nlev; ++l)
          if (H.A[l]) {
            ws_free(c, "sd" + std::to_string(l) + "A:8");
            H.A[l] = nullptr;
          }
    }
    try {
      if (o->mixed == 2) refine_fp32(c, H, F, us, rs, n, S, o, st, rden, sums);
      else pcg(c, H, o->mixed ? &F : nullptr, us, rs, n, S, o, st, rden, sums);
    } catch (...) {
      finish();
      throw;
    }
    return finish();
  }
  if (!o->mixed) {
    for (int cyc = 0; cyc < o->max_cycles; ++cyc) {
      vcycle<double>(c, H, 0, S.ng, S);
      const double rr = read_scalar<double>(c, H.scal + 1) / rden;
      if (st) { st->relres[cyc] = rr; st->cycles = cyc + 1; }
      if (rr <= o->tol) return finish();
    }
    finish();
    throw MmhError(MMH_ENOCONV, "solve: tolerance not reached");
  }
  // mixed defect correction: r = rhs - L u in FP64, an FP32 V-cycle solves L e = r, u += e.
  Hier<float> F;
  build_fp32(c, F, H);
  double* r64 = H.r[0];
  const unsigned nb = stencil_grid(n, 1);
  double* bmax = (double*)ws_get(c, "sbmax", sizeof(double) * nb);
  for (int cyc = 0; cyc < o->max_cycles; ++cyc) {
    op_resid<double>(c, r64, bmax, us, rs, H.G[0], n, 1.0);
    k_convert<double, float><<<nblk(n3), 256, 0, c->stream>>>(F.rhs[0], r64, n3);
    fill<float>(c, F.u[0], 0.f, n3);
    LAUNCHCHK();
    vcycle_mixed(c, F, H, 0, S.ng, S);
    k_add_f2d<<<nblk(n3), 256, 0, c->stream>>>(us, F.u[0], n3);
    LAUNCHCHK();
    const unsigned nr = op_resid<double>(c, (double*)nullptr, bmax, us, rs, H.G[0], n, 1.0);
    reduce<double>(c, bmax, nr, 1, sums + 3);
    const double rr = read_scalar<double>(c, sums + 3) / rden;
    if (st) { st->relres[cyc] = rr; st->cycles = cyc + 1; }
    if (rr <= o->tol) {
      // remove the mean as the FP64 path's top-level zerosum does
      reduce<double>(c, us, n3, 0, sums + 2);
      k_zerosum_guard<double><<<nblk(n3), 256, 0, c->stream>>>(us, sums + 2, H.scal, n3, 0, nullptr);
      LAUNCHCHK();
      return finish();
    }
  }
  finish();
  throw MmhError(MMH_ENOCONV, "solve: tolerance not reached");
}

// ------------------------------------------------------------------------------------------------
// Binned power spectrum and the adjoint of the first outer iteration's mesh solve.
std::vector<double> log_edges(int n, int nb) {
  std::vector<double> e(nb + 1);
  const double k0 = 2.0 * M_PI / n, k1 = M_PI;
  for (int b = 0; b <= nb; ++b) e[b] = k0 * std::exp(std::log(k1 / k0) * (double)b / (double)nb);
  e[0] = k0;
  e[nb] = k1;
  return e;
}

// Forward NR real FFT of f into (sd, sq) and per-bin sums of mult*|F|^2 (S) and mode counts (N).
void spectrum_bins(mmh_ctx* c, const double* f, double* sd, double* sq, int n, int nb, const double* edges_dev,
                   std::vector<double>& S, std::vector<double>& N) {
  const i64 n3 = cube(n);
  HIPCHK(hipMemcpyAsync(sd, f, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  fill<double>(c, sq, 0.0, 2 * (i64)n * n);
  rlft3_dev<double>(c, sd, sq, n, 1);
  const i64 cnt = (i64)(n / 2 + 1) * n * n;
  const unsigned nblk_ = nblk(cnt);
  double* part = (double*)ws_get(c, "pw_part", sizeof(double) * 2 * nb * (size_t)nblk_);
  k_bin_power<<<nblk_, 256, 0, c->stream>>>(part, sd, sq, n, edges_dev, nb);
  LAUNCHCHK();
  std::vector<double> h(2 * (size_t)nb * nblk_);
  HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  S.assign(nb, 0.0);
  N.assign(nb, 0.0);
  for (unsigned b = 0; b < nblk_; ++b)
    for (int q = 0; q < nb; ++q) {
      S[q] += h[(size_t)b * 2 * nb + q];
      N[q] += h[(size_t)b * 2 * nb + nb + q];
    }
}

double* upload_edges(mmh_ctx* c, int n, int nb) {
  const std::vector<double> e = log_edges(n, nb);
  double* d = (double*)ws_get(c, "pw_edges", sizeof(double) * (nb + 1));
  HIPCHK(hipMemcpyAsync(d, e.data(), sizeof(double) * e.size(), hipMemcpyHostToDevice, c->stream));
  return d;
}

void power_dev(mmh_ctx* c, const double* f, int n, int nb, double* pk, double* nmodes) {
  require(is_pow2(n) && n >= 4 && n <= 1024, "power: n must be a power of two in [4, 1024]");
  require(nb >= 1 && nb <= 64, "power: nbins in [1, 64]");
  const i64 n3 = cube(n);
  double* sd = wsT<double>(c, "pw_sd", n3);
  double* sq = wsT<double>(c, "pw_sq", 2 * (i64)n * n);
  double* e = upload_edges(c, n, nb);
  std::vector<double> S, N;
  spectrum_bins(c, f, sd, sq, n, nb, e, S, N);
  for (int b = 0; b < nb; ++b) {
    if (pk) pk[b] = N[b] > 0 ? S[b] / (N[b] * (double)n3) : 0.0;
    if (nmodes) nmodes[b] = N[b];
  }
}

void recon_power_dev(mmh_ctx* c, const double* rho, const double* defin, int n, int nb, const double* w, double tol,
                     double* pk_out, double* grad, double* grad_def, mmh_adjoint_stats* st) {
  require(is_pow2(n) && n >= 16 && n <= 1024, "recon_power: n must be a power of two in [16, 1024]");
  require(nb >= 1 && nb <= 64 && w != nullptr, "recon_power: nbins in [1, 64] and weights required");
  require(tol > 0 && tol < 1, "recon_power: tol in (0, 1)");
  const i64 n3 = cube(n);
  const double pi = 3.14159265358979323846264338328;
  double* u = wsT<double>(c, "ad_u", 5 * n3);
  double* rhs = wsT<double>(c, "ad_rhs", n3);
  double* phi = wsT<double>(c, "ad_phi", n3);
  double* w1 = wsT<double>(c, "ad_w1", n3);
  double* dl = wsT<double>(c, "ad_delta", n3);
  double* sd = wsT<double>(c, "pw_sd", n3);
  double* sq = wsT<double>(c, "pw_sq", 2 * (i64)n * n);
  double* gate = wsT<double>(c, "ad_gate", n3);
  double* det = wsT<double>(c, "ad_det", n3);
  double* sums = (double*)ws_get(c, "ad_sums", sizeof(double) * 8);
  k_to_aos<double><<<nblk(n3), 256, 0, c->stream>>>(u, rho, n3);
  LAUNCHCHK();
  const double* d0 = defin;
  if (!d0) {
    double* z = wsT<double>(c, "ad_def", n3);
    fill<double>(c, z, 0.0, n3);
    d0 = z;
  }
  // forward: rhs (gates recorded), its sum (rgzerosum's), solve, defp = S^2 phi, compression
  // check, delta = D(def + defp), binned power
  std::vector<double> hb;
  limiter_rhs_dev<double>(c, rhs, d0, u, 1.0, 1.0, 5, n, 10.0, hb, gate);
  reduce<double>(c, rhs, n3, 0, sums + 0);  // sum r
  k_det<double><<<nblk(n3), 256, 0, c->stream>>>(det, d0, n, 1.0);
  LAUNCHCHK();
  reduce<double>(c, det, n3, 0, sums + 1);  // sum det (rgzerosum's, unclamped)
  fill<double>(c, phi, 0.0, n3);
  mmh_solve_opts o{};
  o.tol = tol; o.max_cycles = 256; o.schedule = 1; o.nu_pre = 2; o.nu_post = 2; o.nu_coarse = 8; o.ng = 0;
  o.mixed = 0; o.krylov = 1;
  mmh_solve_stats s1;
  solve_dev(c, phi, rhs, d0, u, n, 5, &o, &s1);
  HIPCHK(hipMemcpyAsync(w1, phi, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  nsmooth_dev<double>(c, w1, n, 2);
  const unsigned nbk = nblk(n3);
  double* lim = wsT<double>(c, "ad_lim", n3);
  double* bst = (double*)ws_get(c, "ad_limstat", sizeof(double) * 2 * nbk);
  k_complim<double><<<nbk, 256, 0, c->stream>>>(lim, w1, d0, n, 1.0, 10.0, pi, bst);
  LAUNCHCHK();
  std::vector<double> hc(2 * (size_t)nbk);
  HIPCHK(hipMemcpyAsync(hc.data(), bst, sizeof(double) * hc.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  long ncomp = 0;
  for (unsigned b = 0; b < nbk; ++b) ncomp += (long)hc[2 * b];
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->cycles_forward = s1.cycles;
    st->relres_forward = s1.cycles > 0 ? s1.relres[s1.cycles - 1] : 0.0;
    st->ncompress = (int)ncomp;
  }
  if (ncomp > 0)
    throw MmhError(MMH_EINVAL, "recon_power: the compression limiter fires at this input; the gradient is defined "
                               "only where it is inactive");
  if (defin) k_add<double><<<nblk(n3), 256, 0, c->stream>>>(w1, defin, n3);  // def + defp
  spectral_out<double>(c, w1, dl, n, 1);
  double* e = upload_edges(c, n, nb);
  std::vector<double> S, N;
  spectrum_bins(c, dl, sd, sq, n, nb, e, S, N);
  double loss = 0.0;
  std::vector<double> sym(nb, 0.0);
  for (int b = 0; b < nb; ++b) {
    const double p = N[b] > 0 ? S[b] / (N[b] * (double)n3) : 0.0;
    if (pk_out) pk_out[b] = p;
    loss += w[b] * p;
    sym[b] = N[b] > 0 ? 2.0 * w[b] / N[b] : 0.0;  // d loss / d delta = 2 N^3 IFFT(W F), W = w/(N_b N^3)
  }
  if (st) st->loss = loss;
  if (!grad && !grad_def) return;
  // backward: g_delta = IFFT(sym FFT delta) with sym = 2 w_b / N_b; g_defp = D g_delta (also the
  // direct part of g_def, delta being D(def + defp)); g_phi = S^2 g_defp, mean removed
  double* symd = (double*)ws_get(c, "ad_sym", sizeof(double) * nb);
  HIPCHK(hipMemcpyAsync(symd, sym.data(), sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
  const i64 cnt = (i64)(n / 2 + 1) * n * n;
  k_bin_filter<<<nblk(cnt), 256, 0, c->stream>>>(sd, sq, n, e, nb, symd);
  LAUNCHCHK();
  rlft3_dev<double>(c, sd, sq, n, -1);
  k_scale2<double><<<nblk(n3), 256, 0, c->stream>>>(w1, sd, n, n3);
  LAUNCHCHK();
  spectral_out<double>(c, w1, rhs, n, 1);
  if (grad_def) HIPCHK(hipMemcpyAsync(grad_def, rhs, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  nsmooth_dev<double>(c, rhs, n, 2);
  reduce<double>(c, rhs, n3, 0, sums + 2);
  k_sub_mean<<<nblk(n3), 256, 0, c->stream>>>(rhs, sums + 2, n3);
  LAUNCHCHK();
  // adjoint solve L lambda = P g_phi (L symmetric; the projected right-hand side is mean-free, so
  // the solve's own rgzerosum leaves it unchanged)
  double* lam = dl;
  fill<double>(c, lam, 0.0, n3);
  mmh_solve_stats s2;
  solve_dev(c, lam, rhs, d0, u, n, 5, &o, &s2);
  if (st) {
    st->cycles_adjoint = s2.cycles;
    st->relres_adjoint = s2.cycles > 0 ? s2.relres[s2.cycles - 1] : 0.0;
  }
  // rgzerosum: s = sum r / sum det; ld = sum lambda det
  k_mul<<<nblk(n3), 256, 0, c->stream>>>(w1, lam, det, n3);
  LAUNCHCHK();
  reduce<double>(c, w1, n3, 0, sums + 3);
  double hs[4];
  HIPCHK(hipMemcpyAsync(hs, sums, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double sfac[2] = {hs[0] / hs[1], hs[3] / hs[1]};
  HIPCHK(hipMemcpyAsync(sums + 4, sfac, sizeof(sfac), hipMemcpyHostToDevice, c->stream));
  // through the metric (and rgzerosum's det): g_rhobar, S = dJ/dA
  double* Sb = wsT<double>(c, "ad_S", 6 * n3);
  k_adj_metric<<<nblk(n3), 256, 0, c->stream>>>(w1, Sb, lam, phi, d0, rho, n, sums + 4);
  LAUNCHCHK();
  if (grad_def) {
    k_adj_hess<<<nblk(n3), 256, 0, c->stream>>>(grad_def, Sb, n);
    LAUNCHCHK();
  }
  if (!grad) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return;
  }
  k_box27_adj<<<nblk(n3), 256, 0, c->stream>>>(grad, w1, n);
  LAUNCHCHK();
  // through the right-hand side: g_r = Q^T lambda = lambda - ld / sum det; then nsmooth, the
  // eflux gates, nsmooth^3, the clip mask and 1/30
  HIPCHK(hipMemcpyAsync(w1, lam, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
  k_sub_val<<<nblk(n3), 256, 0, c->stream>>>(w1, sums + 5, n3);
  LAUNCHCHK();
  nsmooth_dev<double>(c, w1, n, 1);
  k_mul<<<nblk(n3), 256, 0, c->stream>>>(w1, w1, gate, n3);
  LAUNCHCHK();
  nsmooth_dev<double>(c, w1, n, 3);
  k_rhs_adj<<<nblk(n3), 256, 0, c->stream>>>(grad, w1, rho, n3, 1.0);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Times ns sweeps of the multigrid smoother on copies of phi, rhs and G (column-major on entry)
// in the layout the V-cycle uses; phi receives the result.
template <class T> void time_sweeps(mmh_ctx* c, T* phi, const T* rhs, const T* G, int n, T dx, int ns, float* ms) {
  const i64 n3 = cube(n);
  T* ps = wsT<T>(c, "ts_p", n3);
  T* rs = wsT<T>(c, "ts_r", n3);
  T* gs = wsT<T>(c, "ts_g", 6 * n3);
  const int sl = c->sweep_impl != 1;
  const int wave_saved = c->wave;
  if (c->sweep_impl == 2) c->wave = 0;
  relayout(c, ps, phi, n, 1, 1);
  relayout(c, rs, rhs, n, 1, 1);
  relayout(c, gs, G, n, 6, 1);
  if (!sl) {
    HIPCHK(hipMemcpyAsync(ps, phi, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(rs, rhs, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(gs, G, sizeof(T) * 6 * n3, hipMemcpyDeviceToDevice, c->stream));
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  sweeps(c, ps, rs, gs, n, dx, ns, sl);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t / std::max(1, ns);
  c->wave = wave_saved;
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  if (sl) relayout(c, phi, ps, n, 1, 0);
  else HIPCHK(hipMemcpyAsync(phi, ps, sizeof(T) * n3, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// FP32 sweeps reading 16-bit (bfloat16) coefficients, the finest-level smoother of the mixed
// Krylov solve: G (6 n^3 FP64, column-major) is rounded to bfloat16 as that solve rounds it.
void time_sweeps_fb(mmh_ctx* c, float* phi, const float* rhs, const double* G, int n, float dx, int ns, float* ms) {
  const i64 n3 = cube(n);
  float* ps = wsT<float>(c, "tsb_p", n3);
  float* rs = wsT<float>(c, "tsb_r", n3);
  bf16* gb = wsT<bf16>(c, "tsb_gb", 6 * n3);
  {
    double* gd = wsT<double>(c, "tsb_gd", 6 * n3);
    relayout(c, gd, G, n, 6, 1);
    convert_bf16(c, gb, gd, 6 * n3);
    ws_free(c, "tsb_gd:8");
  }
  relayout(c, ps, phi, n, 1, 1);
  relayout(c, rs, rhs, n, 1, 1);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  c->slab_now = c->sweep_slabs;  // as the mixed solve's finest level runs them
  sweeps<float, bf16>(c, ps, rs, gb, n, dx, ns, 1);
  c->slab_now = 0;
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t / std::max(1, ns);
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  relayout(c, phi, ps, n, 1, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Host-pointer wrapper: copies inputs to context buffers, runs f, copies outputs back.
struct HostBuf {
  mmh_ctx* c;
  void* d;
  void* h;
  size_t bytes;
  bool out;
};

template <class F> int guarded(mmh_ctx* c, F&& f) {
  if (!c) return MMH_EINVAL;
  try {
    HIPCHK(hipSetDevice(c->device));
    f();
    c->err.clear();
    return MMH_OK;
  } catch (const MmhError& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return MMH_EDEVICE;
  }
}

template <class T> T* up(mmh_ctx* c, const std::string& key, const T* h, i64 count) {
  T* d = wsT<T>(c, "host_" + key, count);
  if (h) HIPCHK(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
  return d;
}
template <class T> void down(mmh_ctx* c, T* h, const T* d, i64 count) {
  HIPCHK(hipMemcpyAsync(h, d, sizeof(T) * count, hipMemcpyDeviceToHost, c->stream));
}

#include "batch.hpp"

std::mutex g_default_mu;
mmh_ctx* g_default = nullptr;

}  // namespace

// ================================================================================================
extern "C" {

int mmh_create(mmh_ctx** out, int device) {
  if (!out) return MMH_EINVAL;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0 || device < 0 || device >= nd) return MMH_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return MMH_EDEVICE;
  scratch_setup(device);
  mmh_ctx* c = new mmh_ctx();
  c->device = device;
  *out = c;
  return MMH_OK;
}

int mmh_release_workspace(mmh_ctx* c) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& kv : c->ws) HIPCHK(hipFree(kv.second.first));
    c->ws.clear();
    c->uploaded.clear();
  });
}

int mmh_destroy(mmh_ctx* c) {
  if (!c) return MMH_EINVAL;
  int rc = mmh_release_workspace(c);
  delete c;
  return rc;
}

const char* mmh_error_string(mmh_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mmh_set_option(mmh_ctx* c, const char* key, int value) {
  return guarded(c, [&] {
    require(key != nullptr, "set_option: null key");
    const std::string k(key);
    if (k == "sweep_impl") {
      require(value >= 0 && value <= 2, "set_option: sweep_impl is 0, 1 or 2");
      c->sweep_impl = value;
    } else if (k == "march_op") {
      c->march_op = value ? 1 : 0;
    } else if (k == "march_tile") {
      require(value >= 0 && value <= 2, "set_option: march_tile is 0, 1 or 2");
      c->march_tile = value;
    } else if (k == "galerkin") {
      c->galerkin = value ? 1 : 0;
    } else if (k == "nu_top") {
      require(value >= 0 && value <= 16, "set_option: nu_top is 0 to 16");
      c->nu_top = value;
    } else if (k == "coarse_fused") {
      require(value == 0 || value == 4 || value == 8 || value == 16 || value == 32,
              "set_option: coarse_fused is 0 (off) or the largest fused level size, 4 to 32");
      c->coarse_fused = value;
    } else if (k == "zero_start") {
      c->zero_start = value ? 1 : 0;
    } else if (k == "colour_col_min_n") {
      require(value >= 8, "set_option: colour_col_min_n >= 8");
      c->colour_col_min_n = value;
    } else if (k == "colour_col") {
      require(value == 1 || value == 2 || value == 4 || value == 8 || value == 16,
              "set_option: colour_col is 1, 2, 4, 8 or 16");
      c->colour_col = value;
    } else if (k == "batch_streams") {
      c->batch_streams = value ? 1 : 0;
    } else if (k == "coarse_fused_ref") {
      require(value == 0 || value == 4 || value == 8 || value == 16 || value == 32,
              "set_option: coarse_fused_ref is 0 (off) or the largest fused level size, 4 to 32");
      c->coarse_fused_ref = value;
    } else if (k == "pcg_graph") {
      c->pcg_graph = value ? 1 : 0;
    } else if (k == "mixed_delta") {
      c->mixed_delta = value ? 1 : 0;
    } else if (k == "wave") {
      c->wave = value ? 1 : 0;
    } else if (k == "wave_p") {
      require(value == 1 || value == 2 || value == 4, "set_option: wave_p is 1, 2 or 4");
      c->wave_p = value;
    } else if (k == "wave_min_n") {
      require(value >= 16, "set_option: wave_min_n >= 16");
      c->wave_min_n = value;
    } else if (k == "batch_graph") {
      c->batch_graph = value ? 1 : 0;
    } else if (k == "sweep_slabs") {
      require(value == 0 || value == 2 || value == 4 || value == 8 || value == 16 || value == 32,
              "set_option: sweep_slabs is 0 (off) or 2, 4, 8, 16 or 32");
      c->sweep_slabs = value;
    } else if (k == "coarse_lds") {
      c->coarse_lds = value ? 1 : 0;
    } else if (k == "dots32") {
      c->dots32 = value ? 1 : 0;
    } else if (k == "batch_coarse_fused") {
      require(value == 0 || value == 4 || value == 8 || value == 16 || value == 32,
              "set_option: batch_coarse_fused is 0 (off) or the largest level size run one workgroup per grid, 4 to 32");
      c->batch_coarse_fused = value;
    } else if (k == "batch_persistent") {
      require(value == 0 || value == 8 || value == 16 || value == 32 || value == 64,
              "set_option: batch_persistent is 0 (off) or the largest grid size solved that way, 8 to 64");
      c->batch_persistent = value;
    } else if (k == "precond_bf16") {
      c->precond_bf16 = value ? 1 : 0;
    } else if (k == "mesh_solve") {
      require(value >= 0 && value <= 2, "set_option: mesh_solve is 0, 1 or 2");
      c->mesh_solve = value;
    } else if (k == "mesh_tol") {
      require(value >= 1 && value <= 14, "set_option: mesh_tol is 1 to 14");
      c->mesh_tol = value;
    } else if (k == "batch_autotune") {
      c->batch_autotune = value ? 1 : 0;
      if (!value) c->tune.clear();
    } else {
      throw MmhError(MMH_EINVAL, std::string("set_option: unknown key ") + key);
    }
  });
}

int mmh_get_option(mmh_ctx* c, const char* key, int* value) {
  return guarded(c, [&] {
    require(key != nullptr && value != nullptr, "get_option: null argument");
    const std::string k(key);
    const std::map<std::string, int> v = {
        {"sweep_impl", c->sweep_impl}, {"march_op", c->march_op}, {"march_tile", c->march_tile},
        {"galerkin", c->galerkin}, {"nu_top", c->nu_top}, {"coarse_fused", c->coarse_fused},
        {"zero_start", c->zero_start}, {"colour_col_min_n", c->colour_col_min_n}, {"colour_col", c->colour_col},
        {"batch_streams", c->batch_streams}, {"coarse_fused_ref", c->coarse_fused_ref}, {"pcg_graph", c->pcg_graph},
        {"mixed_delta", c->mixed_delta}, {"wave", c->wave}, {"wave_p", c->wave_p}, {"wave_min_n", c->wave_min_n},
        {"batch_graph", c->batch_graph}, {"batch_autotune", c->batch_autotune}, {"mesh_solve", c->mesh_solve},
        {"mesh_tol", c->mesh_tol}, {"mesh_cycles", c->mesh_cycles}, {"precond_bf16", c->precond_bf16},
        {"batch_persistent", c->batch_persistent}, {"batch_coarse_fused", c->batch_coarse_fused},
        {"coarse_lds", c->coarse_lds}, {"sweep_slabs", c->sweep_slabs},
        {"dots32", c->dots32}};
    auto it = v.find(k);
    if (it == v.end()) throw MmhError(MMH_EINVAL, std::string("get_option: unknown key ") + key);
    *value = it->second;
  });
}

int mmh_set_stream(mmh_ctx* c, void* s) {
  return guarded(c, [&] { c->stream = (hipStream_t)s; });
}

mmh_ctx* mmh_default_ctx(void) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    if (mmh_create(&g_default, 0) != MMH_OK) g_default = nullptr;
  }
  return g_default;
}

#define DEF_DEV(T, S)                                                                                              \
  int mmh_relax_##S(mmh_ctx* c, T* arr, T* resid, const T* rhs, const T* dp, const T* u, int nu, T* rinf, T dx,  \
                    int n, int nrelax, int iopt) {                                                                 \
    return guarded(c, [&] { relax_api<T>(c, arr, resid, rhs, dp, u, nu, rinf, dx, n, nrelax, iopt); });             \
  }                                                                                                                \
  int mmh_multigrid_##S(mmh_ctx* c, T* u, T* rhs, const T* dp, const T* ub, T dx, int n, int nu, int iopt, int ng, \
                        mmh_mg_stats* st) {                                                                        \
    return guarded(c, [&] { multigrid_dev<T>(c, u, rhs, dp, ub, dx, n, nu, iopt, ng, st); });                       \
  }                                                                                                                \
  int mmh_calcdefp_##S(mmh_ctx* c, T* defp, T* tmp, T* tmp2, const T* def, const T* u, T dtold, T dtaumesh,        \
                       int nf, int n, T cmp, int ng, mmh_lim_stats* st) {                                          \
    return guarded(c, [&] {                                                                                        \
      calcdefp_dev<T>(c, defp, tmp, tmp2, def, u, dtold, dtaumesh, nf, n, cmp, ng, st);                            \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_relaxing_##S(mmh_ctx* c, T* u, const T* def, const T* defp, T dt, int nf, int n) {                       \
    return guarded(c, [&] { relaxing_dev<T>(c, u, def, defp, dt, nf, n); });                                        \
  }                                                                                                                \
  int mmh_reconstruct_##S(mmh_ctx* c, T* u, T* def, T* defp, int n, int nf, int niter, T dtold, T dtaumesh, T dt,   \
                          T cmp, mmh_iter_stats* st) {                                                             \
    return guarded(c, [&] { reconstruct_dev<T>(c, u, def, defp, n, nf, niter, dtold, dtaumesh, dt, cmp, st); });   \
  }                                                                                                                \
  int mmh_gradient_##S(mmh_ctx* c, const T* phi, T* psi, int n) {                                                  \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && n >= 4, "gradient: n must be a power of two >= 4");                                    \
      for (int d = 0; d < 3; ++d) spectral_out<T>(c, phi, psi + (i64)d * cube(n), n, 2 + d);                       \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_delta_##S(mmh_ctx* c, const T* phi, T* delta, int n) {                                                   \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && n >= 4, "delta: n must be a power of two >= 4");                                       \
      spectral_out<T>(c, phi, delta, n, 1);                                                                        \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_nsmooth_##S(mmh_ctx* c, T* a, int n) {                                                                   \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n), "nsmooth: n must be a power of two >= 4");                                               \
      nsmooth_dev<T>(c, a, n, 1);                                                                                  \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_restrict_##S(mmh_ctx* c, T* half, const T* full, int n, int stride) {                                    \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && stride >= 1, "restrict: bad size");                                                    \
      k_restrict27<T><<<nblk(cube(n / 2)), 256, 0, c->stream>>>(half, full, n, stride);                            \
      LAUNCHCHK();                                                                                                 \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_inject_##S(mmh_ctx* c, T* u, const T* uh, int n) {                                                       \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n), "inject: bad size");                                                                     \
      k_prolong<T, 0><<<nblk(cube(n)), 256, 0, c->stream>>>(u, uh, n, 1);                                             \
      LAUNCHCHK();                                                                                                 \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_metric_##S(mmh_ctx* c, T* G, const T* def, const T* u, int nu, int n, T dx, int lsq) {                   \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && nu >= 1, "metric: bad size");                                                          \
      k_metric<T, 0><<<nblk(cube(n)), 256, 0, c->stream>>>(G, def, u, nu, n, dx, lsq);                               \
      LAUNCHCHK();                                                                                                 \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_time_sweeps_##S(mmh_ctx* c, T* phi, const T* rhs, const T* G, int n, T dx, int ns, float* ms) {          \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && ns >= 1 && ms, "time_sweeps: bad arguments");                                          \
      time_sweeps<T>(c, phi, rhs, G, n, dx, ns, ms);                                                               \
    });                                                                                                            \
  }

DEF_DEV(double, d)

int mmh_time_sweeps_fb(mmh_ctx* c, float* phi, const float* rhs, const double* G, int n, float dx, int ns, float* ms) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 4 && ns >= 1 && phi && rhs && G && ms, "time_sweeps_fb: bad arguments");
    time_sweeps_fb(c, phi, rhs, G, n, dx, ns, ms);
  });
}

/* multigrid without allocation: the V-cycle of mmh_multigrid_d with every workspace carved from
 * the device's static scratch array instead of the context's cached buffers. */
int mmh_multigrid_noalloc_d(mmh_ctx* c, double* u, double* rhs, const double* dp, const double* ub, double dx, int n,
                            int nu, int iopt, int ng, mmh_mg_stats* st) {
  return guarded(c, [&] {
    void* base = scratch_ptr(c->device);
    if (!base) {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      const hipError_t e = g_scratch_rc.count(c->device) ? g_scratch_rc[c->device] : hipErrorNotInitialized;
      throw MmhError(MMH_EDEVICE, std::string("multigrid_noalloc: no scratch address for this device: ") +
                                      hipGetErrorString(e));
    }
    require(is_pow2(n) && n <= 64, "multigrid_noalloc: n must be a power of two in [4, 64]");
    Arena A;
    A.base = (unsigned char*)base;
    A.cap = MMH_SCRATCH_BYTES;
    c->arena = &A;
    try {
      multigrid_dev<double>(c, u, rhs, dp, ub, dx, n, nu, iopt, ng, st);
    } catch (...) {
      c->arena = nullptr;
      throw;
    }
    c->arena = nullptr;
  });
}
DEF_DEV(float, f)

int mmh_solve_d(mmh_ctx* c, double* u, double* rhs, const double* defpot, const double* ubig, int n, int nu,
                const mmh_solve_opts* opts, mmh_solve_stats* st) {
  return guarded(c, [&] { solve_dev(c, u, rhs, defpot, ubig, n, nu, opts, st); });
}

int mmh_limiter_rhs_d(mmh_ctx* c, double* tmp, const double* def, const double* u, double dtold, double dtaumesh, int nf,
                      int n, double compressmax) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 16 && n <= 1024 && nf >= 4, "limiter_rhs: bad size");
    require(compressmax >= 0, "limiter_rhs: compressmax < 0");
    std::vector<double> hb;
    limiter_rhs_dev<double>(c, tmp, def, u, dtold < 1.e-8 ? 1.0 : dtold, dtaumesh, nf, n, compressmax, hb);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_residual_d(mmh_ctx* c, double* r, const double* phi, const double* rhs, const double* def, const double* u,
                   int nu, int n, double dx) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 4 && n <= 1024 && nu >= 1, "residual: bad size");
    const i64 n3 = cube(n);
    double* G = wsT<double>(c, "res_G", 6 * n3);
    k_metric<double, 0><<<nblk(n3), 256, 0, c->stream>>>(G, def, u, nu, n, dx, 0);
    const unsigned nb = stencil_grid(n, 0);
    double* bm = (double*)ws_get(c, "res_bm", sizeof(double) * nb);
    k_resid<double, 0><<<nb, RB, 0, c->stream>>>(r, bm, phi, rhs, G, n, 1.0 / (dx * dx));
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_galerkin_d(mmh_ctx* c, double* A_out, const double* defpot, const double* ubig, int nu, int n, double dx,
                   int level) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 8 && n <= 1024 && nu >= 1, "galerkin: bad size");
    require(level >= 1 && (n >> level) >= 4, "galerkin: level out of range");
    const i64 n3 = cube(n);
    double* us = wsT<double>(c, "gk_us", n3);
    double* rs = wsT<double>(c, "gk_rs", n3);
    Hier<double> H;
    build_hier(c, H, "gk", us, rs, defpot, ubig, nu, dx, n, 0);
    build_galerkin(c, H, "gk");
    const int m = n >> level;
    relayout(c, A_out, H.A[level], m, 27, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_solve_batch_d(mmh_ctx* c, int count, const int* n, double* const* u, double* const* rhs,
                      double* const* defpot, double* const* ubig, int nu, const mmh_solve_opts* opts,
                      mmh_batch_stats* st) {
  return guarded(c, [&] { solve_batch_dev(c, count, n, u, rhs, defpot, ubig, nu, opts, st); });
}

int mmh_particle_density_f(mmh_ctx* c, const float* xp, long long np, int ng, float nf, float* den, int* wp_out) {
  return guarded(c, [&] {
    require(xp && den && np >= 1 && np < (1LL << 31), "particle_density: 1 <= np < 2^31 particles required");
    require(ng >= 2 && ng <= 1024 && nf > 0, "particle_density: 2 <= ng <= 1024, nf > 0");
    const i64 ncell = cube(ng);
    float* xs = wsT<float>(c, "pd_xs", 3 * np);
    int* cell = (int*)ws_get(c, "pd_cell", sizeof(int) * np);
    int* idx = (int*)ws_get(c, "pd_idx", sizeof(int) * np);
    int* key = (int*)ws_get(c, "pd_key", sizeof(int) * np);
    int* order = (int*)ws_get(c, "pd_order", sizeof(int) * np);
    int* start = (int*)ws_get(c, "pd_start", sizeof(int) * ncell);
    int* end = (int*)ws_get(c, "pd_end", sizeof(int) * ncell);
    int* near = (int*)ws_get(c, "pd_near", sizeof(int) * ncell);
    int* wp = (int*)ws_get(c, "pd_wp", sizeof(int) * np);
    k_pd_cells<<<nblk(np), 256, 0, c->stream>>>(xs, cell, idx, xp, np, ng, nf);
    LAUNCHCHK();
    // particles sorted by cell, stably, so each cell lists its particles by increasing index
    int bits = 1;
    while ((1LL << bits) < ncell) ++bits;
    size_t tmp_bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, cell, key, idx, order, (int)np, 0, bits, c->stream));
    void* tmp = ws_get(c, "pd_sort_tmp", tmp_bytes + 16);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, cell, key, idx, order, (int)np, 0, bits, c->stream));
    k_fill<int><<<nblk(ncell), 256, 0, c->stream>>>(start, 0, ncell);
    k_fill<int><<<nblk(ncell), 256, 0, c->stream>>>(end, 0, ncell);
    k_fill<int><<<nblk(np), 256, 0, c->stream>>>(wp, 1, np);
    k_pd_ranges<<<nblk(np), 256, 0, c->stream>>>(start, end, key, np);
    k_pd_nearest<<<nblk(ncell), 256, 0, c->stream>>>(near, wp, start, end, order, xs, ng, ng / 2 + 1);
    k_pd_density<<<nblk(ncell), 256, 0, c->stream>>>(den, near, wp, start, end, order, ncell);
    LAUNCHCHK();
    double* sum = wsT<double>(c, "pd_sum", 1);
    reduce<float>(c, den, ncell, 0, sum);
    const float fng = (float)ng;
    k_pd_normalise<<<nblk(ncell), 256, 0, c->stream>>>(den, sum, fng * fng * fng, ncell);
    LAUNCHCHK();
    if (wp_out) HIPCHK(hipMemcpyAsync(wp_out, wp, sizeof(int) * np, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_coarsen_f(mmh_ctx* c, const float* den, int n, int levels, float* out) {
  return guarded(c, [&] {
    require(den && out && is_pow2(n) && levels >= 1 && (n >> levels) >= 1, "coarsen: n a power of two, 1 <= levels");
    const float* src = den;
    int m = n;
    for (int l = 0; l < levels; ++l) {
      const int h = m / 2;
      float* dst = l == levels - 1 ? out : wsT<float>(c, "crs" + std::to_string(l % 2), cube(h));
      k_coarsen2<<<nblk(cube(h)), 256, 0, c->stream>>>(dst, src, h);
      LAUNCHCHK();
      src = dst;
      m = h;
    }
    double* sum = wsT<double>(c, "crs_sum", 1);
    reduce<float>(c, out, cube(m), 0, sum);
    k_scale_mean1<<<nblk(cube(m)), 256, 0, c->stream>>>(out, sum, cube(m));
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_column_density_f(mmh_ctx* c, const float* den, const float* def, int ng, int nplot, float* proj) {
  return guarded(c, [&] {
    require(den && def && proj && ng >= 2 && nplot >= 1 && nplot <= ng, "column_density: bad arguments");
    const i64 n3 = cube(ng);
    float* d = wsT<float>(c, "cd_den", n3);
    HIPCHK(hipMemcpyAsync(d, den, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->stream));
    double* sum = wsT<double>(c, "cd_sum", 1);
    reduce<float>(c, d, n3, 0, sum);
    k_scale_mean1<<<nblk(n3), 256, 0, c->stream>>>(d, sum, n3);
    k_column_density<<<nblk((i64)ng * ng), 256, 0, c->stream>>>(proj, d, def, ng, nplot);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_gridfigure_f(mmh_ctx* c, const float* def, int nf, const float* den, int ng, int nplot, float* pos,
                     float* proj, float* dsp_slab) {
  return guarded(c, [&] {
    require(def && den && pos && proj && nf >= 4 && ng >= nf && ng % nf == 0 && nplot >= 1 &&
                nplot <= ng / 2 - ng / nf,
            "gridfigure: nf >= 4, ng a multiple of nf, 1 <= nplot <= ng/2 - ng/nf");
    const int ratio = ng / nf, E = nf + 1;
    const i64 e3 = (i64)E * E * E, n3 = cube(ng);
    float* t = wsT<float>(c, "gf_t", 3 * e3);
    fill<float>(c, t, 0.f, 3 * e3);
    k_grid_corner_dsp<<<nblk(cube(nf)), 256, 0, c->stream>>>(t, def, nf, ratio);
    if (dsp_slab) k_grid_dsp_slab<<<nblk((i64)E * E * 4), 256, 0, c->stream>>>(dsp_slab, t, nf);
    k_grid_ghosts<<<nblk((i64)E * E), 256, 0, c->stream>>>(t, nf);
    k_grid_positions<<<nblk((i64)nf * nf), 256, 0, c->stream>>>(pos, t, nf, ratio, ng);
    LAUNCHCHK();
    float* d = wsT<float>(c, "gf_den", n3);
    HIPCHK(hipMemcpyAsync(d, den, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->stream));
    double* sum = wsT<double>(c, "gf_sum", 1);
    reduce<float>(c, d, n3, 0, sum);
    k_scale_mean1<<<nblk(n3), 256, 0, c->stream>>>(d, sum, n3);
    k_grid_column<<<nblk((i64)ng * ng), 256, 0, c->stream>>>(proj, d, t, nf, ratio, ng, nplot);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_power_d(mmh_ctx* c, const double* field, int n, int nbins, double* pk, double* nmodes) {
  return guarded(c, [&] { power_dev(c, field, n, nbins, pk, nmodes); });
}

int mmh_recon_power_grad_d(mmh_ctx* c, const double* rho, int n, int nbins, const double* w, double tol, double* pk,
                           double* grad, mmh_adjoint_stats* st) {
  return guarded(c, [&] { recon_power_dev(c, rho, nullptr, n, nbins, w, tol, pk, grad, nullptr, st); });
}
int mmh_recon_power_grad_mesh_d(mmh_ctx* c, const double* rho, const double* def, int n, int nbins, const double* w,
                                double tol, double* pk, double* grad_rho, double* grad_def, mmh_adjoint_stats* st) {
  return guarded(c, [&] {
    require(def != nullptr, "recon_power_grad_mesh: def required (mmh_recon_power_grad_d is the Cartesian mesh)");
    recon_power_dev(c, rho, def, n, nbins, w, tol, pk, grad_rho, grad_def, st);
  });
}

int mmh_fisher_d(const double* pk, int nsim, int nk, const double* cv, double* cov, double* covnorm, double* fisher,
                 double* error) {
  if (!pk || !cv || nsim < 2 || nk < 1) return MMH_EINVAL;
  auto P = [&](int s, int b) { return pk[(size_t)s * nk + b]; };
  std::vector<double> sums(nk, 0.0);
  for (int b = 0; b < nk; ++b)
    for (int s = 0; s < nsim; ++s) sums[b] += P(s, b);
  std::vector<double> cv_(nk * (size_t)nk), cn(nk * (size_t)nk);
  // calccov: centred series, dot product / (nsim - 1); normalised by the means
  for (int y = 0; y < nk; ++y)
    for (int x = 0; x < nk; ++x) {
      const double m1 = sums[x] / nsim, m2 = sums[y] / nsim;
      double d = 0.0;
      for (int s = 0; s < nsim; ++s) d += (P(s, x) - m1) * (P(s, y) - m2);
      const double c12 = d / (nsim - 1);
      cv_[x + (size_t)nk * y] = c12;
      cn[x + (size_t)nk * y] = c12 / sums[x] / sums[y] * (double)(nsim * nsim);
    }
  if (cov) std::copy(cv_.begin(), cv_.end(), cov);
  if (covnorm) std::copy(cn.begin(), cn.end(), covnorm);
  if (error) {
    // calerror: standard error of the mean
    for (int b = 0; b < nk; ++b) {
      const double m = sums[b] / nsim;
      double q = 0.0;
      for (int s = 0; s < nsim; ++s) q += (P(s, b) - m) * (P(s, b) - m);
      // sqrt(real(n_sim)) is a default-REAL (single precision) square root in calerror
      error[b] = std::sqrt(q / (nsim - 1)) / (double)std::sqrt((float)nsim);
    }
  }
  if (fisher) {
    for (int m = 1; m <= nk; ++m) {
      // Gauss-Jordan without pivoting on the leading m x m block: eliminate below, then above
      std::vector<double> B(m * (size_t)m), I(m * (size_t)m, 0.0);
      for (int jj = 0; jj < m; ++jj)
        for (int ii = 0; ii < m; ++ii) B[ii + (size_t)m * jj] = cn[ii + (size_t)nk * jj];
      for (int ii = 0; ii < m; ++ii) I[ii + (size_t)m * ii] = 1.0;
      auto row_op = [&](int J, int Irow, double E, int c0) {
        for (int col = c0; col < m; ++col) B[J + (size_t)m * col] = B[J + (size_t)m * col] - B[Irow + (size_t)m * col] * E;
        for (int col = 0; col < m; ++col) I[J + (size_t)m * col] = I[J + (size_t)m * col] - I[Irow + (size_t)m * col] * E;
      };
      for (int Ii = 0; Ii < m - 1; ++Ii)
        for (int J = Ii + 1; J < m; ++J) row_op(J, Ii, B[J + (size_t)m * Ii] / B[Ii + (size_t)m * Ii], Ii);
      for (int Ii = m - 1; Ii >= 1; --Ii)
        for (int J = Ii - 1; J >= 0; --J) row_op(J, Ii, B[J + (size_t)m * Ii] / B[Ii + (size_t)m * Ii], 0);
      for (int ii = 0; ii < m; ++ii) {
        const double dg = B[ii + (size_t)m * ii];
        for (int col = 0; col < m; ++col) I[ii + (size_t)m * col] = I[ii + (size_t)m * col] / dg;
      }
      // fishinfo = sum_j (sum_i inv(i,j) c_i) c_j
      double f = 0.0;
      std::vector<double> t(m, 0.0);
      for (int jj = 0; jj < m; ++jj)
        for (int ii = 0; ii < m; ++ii) t[jj] = t[jj] + I[ii + (size_t)m * jj] * cv[ii];
      for (int jj = 0; jj < m; ++jj) f = f + t[jj] * cv[jj];
      fisher[m - 1] = f;
    }
  }
  return MMH_OK;
}

int mmh_rlft3_d(mmh_ctx* c, double* data, double* speq, int n, int isign) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 4 && (isign == 1 || isign == -1), "rlft3: bad arguments");
    rlft3_dev<double>(c, data, speq, n, isign);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_lognormal_d(mmh_ctx* c, double* rho, int n, double sigma, double slope, double ks_frac, uint64_t seed) {
  return guarded(c, [&] { lognormal_dev(c, rho, n, sigma, slope, ks_frac, seed); });
}

// ---- host-pointer entry points -------------------------------------------------------------
#define DEF_HOST(T, S)                                                                                             \
  int mmh_relax_host_##S(mmh_ctx* c, T* arr, T* resid, const T* rhs, const T* dp, const T* u, int nu, T* rinf,     \
                         T dx, int n, int nrelax, int iopt) {                                                      \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && nu >= 1, "relax: bad size");                                                           \
      const i64 n3 = cube(n), h3 = cube(n / 2);                                                                    \
      T* a = up(c, "a", arr, n3);                                                                                  \
      T* r = up(c, "r", resid, h3);                                                                                \
      T* b = up(c, "b", rhs, n3);                                                                                  \
      T* d = up(c, "d", dp, n3);                                                                                   \
      T* w = up(c, "u", u, (i64)nu * n3);                                                                          \
      relax_api<T>(c, a, r, b, d, w, nu, rinf, dx, n, nrelax, iopt);                                               \
      down(c, arr, a, n3);                                                                                         \
      if (resid) down(c, resid, r, h3);                                                                            \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_multigrid_host_##S(mmh_ctx* c, T* u, T* rhs, const T* dp, const T* ub, T dx, int n, int nu, int iopt,    \
                             int ng) {                                                                             \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && nu >= 1, "multigrid: bad size");                                                       \
      const i64 n3 = cube(n);                                                                                      \
      T* a = up(c, "a", u, n3);                                                                                    \
      T* b = up(c, "b", rhs, n3);                                                                                  \
      T* d = up(c, "d", dp, n3);                                                                                   \
      T* w = up(c, "u", ub, (i64)nu * n3);                                                                         \
      multigrid_dev<T>(c, a, b, d, w, dx, n, nu, iopt, ng, nullptr);                                               \
      down(c, u, a, n3);                                                                                           \
      down(c, rhs, b, n3);                                                                                         \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_calcdefp_host_##S(mmh_ctx* c, T* defp, T* tmp, T* tmp2, const T* def, const T* u, T dtold, T dtaumesh,  \
                            int nf, int n, T cmp) {                                                                \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && nf >= 4, "calcdefp: bad size");                                                        \
      const i64 n3 = cube(n), p3 = (i64)(n + 2) * n * n;                                                           \
      T* p = up<T>(c, "a", nullptr, n3);                                                                           \
      T* t = up<T>(c, "b", nullptr, n3);                                                                           \
      T* t2 = up<T>(c, "t2", nullptr, p3);                                                                         \
      T* d = up(c, "d", def, n3);                                                                                  \
      T* w = up(c, "u", u, (i64)nf * n3);                                                                          \
      calcdefp_dev<T>(c, p, t, t2, d, w, dtold, dtaumesh, nf, n, cmp, n, nullptr);                                 \
      down(c, defp, p, n3);                                                                                        \
      down(c, tmp, t, n3);                                                                                         \
      down(c, tmp2, t2, p3);                                                                                       \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }                                                                                                                \
  int mmh_relaxing_host_##S(mmh_ctx* c, T* u, const T* def, const T* defp, T dt, int nf, int n) {                  \
    return guarded(c, [&] {                                                                                        \
      require(is_pow2(n) && nf == 5, "relaxing: bad size");                                                        \
      const i64 n3 = cube(n);                                                                                      \
      T* w = up(c, "u", u, 5 * n3);                                                                                \
      T* d = up(c, "d", def, n3);                                                                                  \
      T* p = up(c, "a", defp, n3);                                                                                 \
      relaxing_dev<T>(c, w, d, p, dt, nf, n);                                                                      \
      down(c, u, w, 5 * n3);                                                                                       \
      HIPCHK(hipStreamSynchronize(c->stream));                                                                     \
    });                                                                                                            \
  }

DEF_HOST(double, d)
DEF_HOST(float, f)

int mmh_gradient_host_f(mmh_ctx* c, const float* phi, float* psi, int n) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 4, "gradient: bad size");
    const i64 n3 = cube(n);
    float* a = up(c, "a", phi, n3);
    float* o = up<float>(c, "g3", nullptr, 3 * n3);
    for (int d = 0; d < 3; ++d) spectral_out<float>(c, a, o + (i64)d * n3, n, 2 + d);
    down(c, psi, o, 3 * n3);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_nsmooth_host_f(mmh_ctx* c, float* a, int n) {
  return guarded(c, [&] {
    require(is_pow2(n), "nsmooth: bad size");
    const i64 n3 = cube(n);
    float* d = up(c, "a", a, n3);
    nsmooth_dev<float>(c, d, n, 1);
    down(c, a, d, n3);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_lognormal_host_d(mmh_ctx* c, double* rho, int n, double sigma, double slope, double ks_frac, uint64_t seed) {
  return guarded(c, [&] {
    require(is_pow2(n), "lognormal: bad size");
    double* d = up<double>(c, "rho", nullptr, cube(n));
    lognormal_dev(c, d, n, sigma, slope, ks_frac, seed);
    down(c, rho, d, cube(n));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int mmh_reconstruct_host_f(mmh_ctx* c, float* u, float* def, float* defp, int n, int nf, int niter, float dtold,
                           float dtaumesh, float dt, float cmp, double* diag) {
  return guarded(c, [&] {
    require(is_pow2(n) && nf == 5 && niter >= 0, "reconstruct: bad size");
    const i64 n3 = cube(n);
    float* w = up(c, "u", u, 5 * n3);
    float* d = up(c, "d", def, n3);
    float* p = up<float>(c, "a", nullptr, n3);
    std::vector<mmh_iter_stats> st(diag ? niter : 0);
    reconstruct_dev<float>(c, w, d, p, n, nf, niter, dtold, dtaumesh, dt, cmp, diag ? st.data() : nullptr);
    down(c, u, w, 5 * n3);
    down(c, def, d, n3);
    down(c, defp, p, n3);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int it = 0; diag && it < niter; ++it) {
      double* o = diag + 7 * (size_t)it;
      o[0] = st[it].defp_sigma; o[1] = st[it].def_sigma; o[2] = st[it].u_sigma;
      o[3] = st[it].u_min; o[4] = st[it].u_max; o[5] = (double)st[it].u_negative; o[6] = st[it].lim.ncompress;
    }
  });
}

int mmh_delta_host_f(mmh_ctx* c, const float* phi, float* delta, int n) {
  return guarded(c, [&] {
    require(is_pow2(n) && n >= 4, "delta: bad size");
    const i64 n3 = cube(n);
    float* a = up(c, "a", phi, n3);
    float* o = up<float>(c, "b", nullptr, n3);
    spectral_out<float>(c, a, o, n, 1);
    down(c, delta, o, n3);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
