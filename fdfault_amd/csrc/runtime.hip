// Device runtime and C ABI of the MI355X time-stepping core (see
// include/fdfault_hip.h).  One HIP stream per handle; f is double-buffered on
// the device so each stage reads f and writes the other buffer.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <ctime>
#include <deque>
#include <memory>
#include <mutex>
#include <random>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/fdfault_hip.h"
#include "kernels.hpp"
#include "setup.hpp"

namespace fdf {

static thread_local std::string g_last_error;

#define HIPCHK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw Error{std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x}; \
  } while (0)

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  // owns its allocation: movable, not copyable
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      free();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  void alloc(size_t count) {
    free();
    n = count;
    if (count) HIPCHK(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const T* h, size_t count) {
    alloc(count);
    if (count) HIPCHK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
  void zero(hipStream_t s) {
    if (n) HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { free(); }
};

struct OutFile {
  FILE* data = nullptr;
  FILE* time = nullptr;
};

// Device-side grid set-up (block.cpp:769-1046, problem.cpp:94-134): the
// problem compiler hands each block's face surfaces here; the grid, metric
// and Jacobian are built by kernels into FP64 device arrays that the solver
// then adopts, and the host reads back only the points it needs.
class DeviceGrid : public GridBuilder {
 public:
  DBuf<double> x, metric, jac;  // [ndim][N], [ndim^2][N], [N]
  DBuf<double> mat;             // heterogeneous material (dt reduction)
  DBuf<int> bad;
  hipStream_t stream = nullptr;
  bool table_set = false;

  explicit DeviceGrid(hipStream_t s) : stream(s) {}

  void build_block(const SetupH& S, const BlockH& B, const std::vector<double>* surf, const int* n0,
                   const int* n1) override {
    if (!x.p) {
      x.alloc((size_t)S.ndim * S.N);
      metric.alloc((size_t)S.ndim * S.ndim * S.N);
      jac.alloc(S.N);
      bad.alloc(1);
      bad.zero(stream);
    }
    if (!table_set) {
      HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_fd), S.fd.c, sizeof(S.fd.c)));
      table_set = true;
    }
    GridBlockDev g{};
    DBuf<double> fb[6];
    for (int f = 0; f < 2 * S.ndim; f++) {
      fb[f].upload(surf[f]);
      g.F[f] = SurfDev{fb[f].p, n0[f], n1[f]};
    }
    for (int d = 0; d < 3; d++) {
      g.nx[d] = B.nx[d];
      g.xm[d] = B.xm[d];
      g.dx[d] = B.dx[d];
    }
    g.ndim = S.ndim;
    g.Ny = S.nx[1];
    g.Nz = S.nx[2];
    g.N = (long long)S.N;
    const long long nb = (long long)B.nx[0] * B.nx[1] * B.nx[2];
    const dim3 grid((unsigned)((nb + 255) / 256));
    hipLaunchKernelGGL(grid_tfi_kernel, grid, dim3(256), 0, stream, g, x.p);
    HIPCHK(hipGetLastError());
    // the metric reads x along the block's own lines only, so one block's
    // grid is complete before its metric is formed (same stream)
    hipLaunchKernelGGL(grid_metric_kernel, grid, dim3(256), 0, stream, g, S.fd.s, x.p, metric.p, jac.p, bad.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));  // the surface buffers are freed on return
  }

  bool jacobians_positive() override {
    int b = 0;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&b, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    return b == 0;
  }

  void gather(const SetupH& S, const std::vector<size_t>& idx, double* ox, double* om, double* oj) override {
    const int n = (int)idx.size();
    if (!n) return;
    std::vector<long long> id(idx.begin(), idx.end());
    DBuf<long long> di;
    di.upload(id);
    DBuf<double> dx, dm, dj;
    dx.alloc((size_t)n * S.ndim);
    dm.alloc((size_t)n * S.ndim * S.ndim);
    dj.alloc((size_t)n);
    hipLaunchKernelGGL(grid_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, di.p, n, S.ndim,
                       (long long)S.N, x.p, metric.p, jac.p, dx.p, dm.p, dj.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(ox, dx.p, dx.n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(om, dm.p, dm.n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(oj, dj.p, dj.n * sizeof(double), hipMemcpyDeviceToHost));
  }

  double block_worst(const SetupH& S, const BlockH& B) override {
    if (S.hetmat && !mat.p) mat.upload(S.mat);
    DBuf<unsigned long long> w;
    w.alloc(1);
    w.zero(stream);
    const long long nb = (long long)B.nx[0] * B.nx[1] * B.nx[2];
    hipLaunchKernelGGL(grid_worst_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream, B.mlb[0],
                       B.mlb[1], B.mlb[2], B.prb[0] - B.mlb[0], B.prb[1] - B.mlb[1], B.prb[2] - B.mlb[2], S.ndim,
                       S.nx[1], S.nx[2], (long long)S.N, B.dx[0], B.dx[1], B.dx[2], B.cs, metric.p,
                       S.hetmat ? mat.p : nullptr, w.p);
    HIPCHK(hipGetLastError());
    unsigned long long bits = 0;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&bits, w.p, sizeof(bits), hipMemcpyDeviceToHost));
    double v;
    memcpy(&v, &bits, sizeof(v));
    return v;
  }

  // block_metric_uniform (setup.cpp) on the device arrays
  bool uniform(const SetupH& S, const BlockH& B, double mdiag[3], const double* metric_p, const double* jac_p) {
    const size_t gc = S.idx((B.mlb[0] + B.prb[0]) / 2, (B.mlb[1] + B.prb[1]) / 2, (B.mlb[2] + B.prb[2]) / 2);
    const double jc = S.jac_at(gc);
    for (int d = 0; d < 3; d++) mdiag[d] = d < S.ndim ? S.metric_at(d * S.ndim + d, gc) : 0.;
    DBuf<int> flag;
    flag.alloc(1);
    flag.zero(stream);
    const long long nb = (long long)B.nx[0] * B.nx[1] * B.nx[2];
    hipLaunchKernelGGL(grid_uniform_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream, B.mlb[0],
                       B.mlb[1], B.mlb[2], B.prb[0] - B.mlb[0], B.prb[1] - B.mlb[1], B.prb[2] - B.mlb[2], S.ndim,
                       S.nx[1], S.nx[2], (long long)S.N, jc, mdiag[0], mdiag[1], mdiag[2], metric_p, jac_p, flag.p);
    HIPCHK(hipGetLastError());
    int b = 0;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&b, flag.p, sizeof(int), hipMemcpyDeviceToHost));
    return b == 0;
  }
};

class SolverBase {
 public:
  virtual ~SolverBase() {}
  virtual void step(int n) = 0;
  virtual void stage(int istep, int st) = 0;
  virtual void read_array(int which, double* h, long n) = 0;
  virtual void write_array(int which, const double* h, long n) = 0;
  virtual void read_field(int field, const int lo[3], const int hi[3], const int st[3], double* h) = 0;
  virtual long read_iface(int iface, int which, double* h, long n) = 0;
  virtual long read_front(int iface, double* h, long n) = 0;
  virtual void finish() = 0;
  virtual void save_state(const char* path) = 0;
  virtual void load_state(const char* path) = 0;
  virtual double bulk_ms(int* n) = 0;
  virtual double fault_ms(int* n) = 0;
  SetupH S;
  fdf_opts opts{};
  int cur_step = 0;
  int ntiled = 0;  // blocks on a rectilinear tiled kernel
  int nrows = 0;   // of which on the row-tile kernel
  int ncurv = 0;   // 2-D blocks on the curvilinear tiled kernel
  int nwide = 0;   // of the tiled and curvilinear blocks, with one descriptor per field
  int nfused = 0;  // interfaces solved inside the merged 2-D stage launch
};

template <typename R>
class Batch;

template <typename R>
class Solver : public SolverBase {
  friend class Batch<R>;

 public:
  explicit Solver(const std::string& path, const fdf_opts& o);
  ~Solver() override;
  void step(int n) override;
  void stage(int istep, int st) override { do_stage(istep, st, false); sync_check(); }
  void read_array(int which, double* h, long n) override;
  void write_array(int which, const double* h, long n) override;
  void read_field(int field, const int lo[3], const int hi[3], const int st[3], double* h) override;
  long read_iface(int iface, int which, double* h, long n) override;
  long read_front(int iface, double* h, long n) override;
  void finish() override;
  void save_state(const char* path) override;
  void load_state(const char* path) override;
  double bulk_ms(int* n) override;
  double fault_ms(int* n) override;

 private:
  int mode_id() const { return S.ndim == 3 ? 2 : (S.mode == 2 ? 0 : 1); }
  void do_stage(int istep, int st, bool full_step);
  // The last stage of a full time step leaves df unwritten when the next
  // step's first stage scales it by A = 0 (every scheme of rk.cpp:42-91):
  // df_stale marks that, and materialize_df() recomputes it, bit-identically,
  // for a caller that reads df (the last stage's input f is still in the
  // other f buffer, df holds the previous stage's, and the interface SAT
  // buffers hold the last stage's terms until the next stage).
  bool df_stale = false;
  int stage_dfmode(int st, bool full_step) const;
  void materialize_df();
  void launch_fault(int i, const FaultStageDev& fs);
  void launch_bulk(int b, const StageDev& sd);
  void launch_tiled(Tiled3D ta, int ty);
  void launch_tiled2(Tiled2D t2, int tw, int par, hipStream_t st);
  Tiled2D stage_tile2(int b, const StageDev& sd) const;
  std::vector<DBuf<R>> fcoefs;  // per block [6][16] external SAT entries
  int tile_ty = 4;
  void launch_plastic(int b);
  void write_outputs(int tstep, hipStream_t st = nullptr);
  // device-side set-up: the builder, and the FP64 grid arrays on the device
  // (the solver's own for FP64, the builder's for FP32)
  std::unique_ptr<DeviceGrid> dgrid;
  const double* metric64 = nullptr;
  const double* jac64 = nullptr;
  // component c of the coordinates over an index box (C order)
  void coords_box(int c, const int lo[3], const int hi[3], const int st[3], std::vector<double>& out);
  void sync_check();

  hipStream_t stream = nullptr;
  static constexpr int NEV = 128;
  hipEvent_t evs[2 * NEV] = {};
  int nev = 0;  // recorded pairs not yet summed
  double bulk_time_ms = 0., fault_time_ms = 0.;
  int bulk_count = 0, fault_count = 0;
  int evkind[NEV] = {};  // 0 bulk launch, 1 the stage's fault launches
  void flush_events();
  GridDev G{};
  DBuf<R> f[2], df, fp, metric, jac, mat;
  DBuf<double> hs, x;
  int cur = 0;
  std::vector<BlockDev> bdev;
  std::vector<bool> tiled_ok;
  std::vector<bool> curv_ok;  // 2-D mode II block on the curvilinear tiled kernel
  std::vector<DBuf<R>> smaps;  // per block and face: the external SAT map of each face point
  // 2-D problems whose blocks all take the same tiled kernel: one launch per
  // stage over every block's tiles (merged 1: rectilinear, 2: curvilinear)
  int merged = 0, cur_stage = 0, mnwork = 0;
  // 3-D problems of two blocks on the tiled kernels: the second block's
  // launch runs on `side`, forked after the stage's fault solve and joined
  // before the next stage, so its workgroups fill the CUs that the first
  // block's tail leaves idle
  bool pair3 = false;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<DBuf<Tiled2D>> mdesc;  // [stage][roundtrip] per-block descriptors
  DBuf<BlockDev> mbd;
  DBuf<int2> mwork;                  // (block, tile)
  // ... and, for one slip-weakening y-interface whose two blocks' tiles line
  // up in x, the same launch with the fault solve in it (bulk_fault2d_batch):
  // work items (block, tile) and (-1 - chunk), one fault descriptor, the
  // chunks (on by default; FDF_FUSE2D=0 turns it off)
  bool fused2d = false;
  int mnfwork = 0;
  DBuf<int2> mfwork;
  DBuf<FaultBatchDev> mfdev;
  DBuf<FaultPair2D> mpairs;
  void launch_fused(const StageDev& sd, const FaultStageDev& fs);
  std::vector<Tiled3D> tiles;
  std::vector<Tiled2D> tiles2;
  int tile2_tw = 64, tile2_cx = 2;  // 2-D tile: y points per workgroup, x planes per chunk
  bool tile2_env = false;            // FDF_TILE2_TW / FDF_TILE2_CX fix the 2-D tile
  struct FaceBufs {
    DBuf<double> nrm, dl;
  };
  std::vector<FaceBufs> fbufs;  // per block*6
  struct IfBufs {
    DBuf<double> nrm, dl1, dl2, load_xy, load_par, lf, pert_xy, pert_par, pf;
    DBuf<double> u, du, v, s, sn, ux, dux, vx, sx, state, dstate, sat, front;
  };
  std::vector<IfBufs> ibufs;
  std::vector<FaultDev> fdev;
  // STZ learned proposal, solve log and counters (shared by all STZ interfaces)
  DBuf<float> stz_model;
  DBuf<double> stz_log;
  DBuf<unsigned long long> stz_cnt;  // [log rows, solves, iterations, proposals used, fallbacks]
  void stz_refresh();
 public:
  void stz_set_model(const float* params, int n);
  void stz_enable_log(long cap);
  long stz_read_log(double* out, long cap);
  void stz_stats(unsigned long long out[4]);
 private:
  DBuf<int> err;
  DBuf<double> stage_buf;
  std::vector<OutFile> outs;
  // Output units without a blocking copy per unit (outputunit.cpp:771-889):
  // on an output step each unit's values are gathered on the device into the
  // next slot of its staging ring, copied to pinned host memory on the copy
  // stream, and written to the unit's file once the copy has landed, in step
  // order (a FIFO over all units).  A slot is reused only after its file
  // write, so the host waits only when a ring is full.
  struct OutStage {
    DBuf<double> dev;        // [nslot][cnt]
    double* host = nullptr;  // pinned, [nslot][cnt]
    size_t cnt = 0;
    int nslot = 0, next = 0;
    std::vector<hipEvent_t> gathered, copied;
    std::vector<char> busy;
  };
  std::vector<OutStage> ostage;
  DBuf<OutGatherDev> ogather;  // per unit: what gather_units copies into its ring
  std::deque<std::pair<int, int>> opend;  // (unit, slot) in step order
  hipStream_t ostream = nullptr;
  void setup_output_staging();
  bool drain_output(bool wait);  // write the oldest pending slot; false if none (or not landed, !wait)
  void drain_outputs();
  DBuf<double>* iface_buf(int iface, int which);
  // graph replay state
  void graph_step();
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  DBuf<int> dstep;
  const int* graph_step_ptr = nullptr;
  bool capturing = false, dstep_valid = false;
  int graph_flip = 0;
};

// row-tile kernel: FP32 4 rows per workgroup (<= 1024 threads, 4 waves per
// SIMD), FP64 2 rows (<= 512 threads, 2 waves per SIMD)
template <typename R, int M>
static const void* fault_kernel_law(int law) {
  switch (law) {
    case 0: return (const void*)fault_kernel<M, R, 0>;
    case 1: return (const void*)fault_kernel<M, R, 1>;
    case 2: return (const void*)fault_kernel<M, R, 2>;
    default: return (const void*)fault_kernel<M, R, 3>;
  }
}

// the fault kernel of a mode (0 2-D mode II, 1 mode III, 2 3-D) and law
template <typename R>
static const void* fault_kernel_for(int mode, int law) {
  if (mode == 0) return fault_kernel_law<R, 0>(law);
  if (mode == 1) return fault_kernel_law<R, 1>(law);
  return fault_kernel_law<R, 2>(law);
}

// The x-chunks of a 2-D y-interface between the blocks of tile geometries A
// (side 1, below) and B (side 2, above), as work items of bulk_fault2d_batch:
// false when the interface cannot be solved inside the stage launch (not
// slip weakening, not along x over both blocks' whole x range, or tiles that
// do not line up).  fd: fault descriptor; dA, dB: block descriptors; yA, yB:
// the y tile of each side's interface row (their tiles join the chunk item).
static bool fault_chunks_2d(const IfaceH& I, const FaultDev& F, const Tiled2D& A, const Tiled2D& B, int tw, int fd,
                            int dA, int dB, std::vector<FaultPair2D>& out, int& yA, int& yB) {
  if (I.law != LAW_SLIPWEAK || I.dir != 1 || F.n1 != 1) return false;
  if (A.lo[0] != B.lo[0] || A.hi[0] != B.hi[0] || A.cx != B.cx || A.nxc != B.nxc || A.cx > tw) return false;
  if (I.org[0] != A.lo[0] || F.n != A.hi[0] - A.lo[0]) return false;
  if (I.org[1] != A.hi[1] - 1 || I.org[1] + 1 != B.lo[1]) return false;
  yA = (I.org[1] - A.lo[1]) / tw;
  yB = (I.org[1] + 1 - B.lo[1]) / tw;
  for (int xc = 0; xc < A.nxc; xc++) {
    const int x0 = A.lo[0] + xc * A.cx;
    out.push_back(FaultPair2D{fd, dA, dB, yA + xc * A.nyt, yB + xc * B.nyt, x0 - I.org[0],
                              std::min(A.cx, A.hi[0] - x0)});
  }
  return true;
}

// Work list of bulk_fault2d_batch: the tile items tl, and npairs chunk
// items (-1 - k, 0) placed first in every XCD's run of the list (workgroup
// b of a grid of G takes item (b % 8)(G / 8) + b / 8 first), so that the
// chunk workgroups, which do the most work, start first and are spread over
// the XCDs; the tile items keep their order.  G <= 0: one workgroup per
// item (G is set to the item count rounded up to 8).
static std::vector<int2> fused_work_list(const std::vector<int2>& tl, int npairs, int& G) {
  const int total = (int)tl.size() + npairs;
  if (G <= 0) G = ((total + 7) / 8) * 8;
  const int per = std::max(1, G / 8);
  std::vector<int2> out((size_t)total);
  std::vector<char> used((size_t)total, 0);
  int k = 0;
  for (int sl = 0; sl < per && k < npairs; sl++)
    for (int x = 0; x < 8 && k < npairs; x++) {
      const long pos = (long)x * per + sl;
      if (pos < total && !used[pos]) {
        out[pos] = make_int2(-1 - k++, 0);
        used[pos] = 1;
      }
    }
  for (int pos = 0; pos < total && k < npairs; pos++)
    if (!used[pos]) {
      out[pos] = make_int2(-1 - k++, 0);
      used[pos] = 1;
    }
  size_t t = 0;
  for (int pos = 0; pos < total; pos++)
    if (!used[pos]) out[pos] = tl[t++];
  return out;
}

// row-kernel variant: 1 = double-buffered LDS tile (FP32; FP64's does not
// fit the LDS), 0 = single buffer
static const int rows_var_default = 1;
// row-tile order: patches of 16 y tiles x 2 x chunks (32 workgroups, one per
// CU of an XCD; measured best of 8x4, 4x8, 16x2, 8x8, 2x16, 32x1)
static const int rows_grp_y = 16, rows_grp_x = 2;

// row-tile workgroup shapes: 0 = FP32 4 rows (<= 1024 threads, 4 waves per
// SIMD) or FP64 2 rows (<= 512 threads, 2 waves per SIMD), LDS rows padded to
// 256 points; 1 = FP32 6 rows of at most 160 points, LDS rows padded to 160;
// 2 = FP32 5 rows of at most 204 points; 3 = FP64 4 rows of at most 128
// points, LDS rows padded to 128; 4 = FP32 sbporder 4 (more than 128
// registers): 6 rows of at most 128 points, <= 768 threads at 3 waves per
// SIMD.  As many rows as fit the thread budget, so that a workgroup fills
// the wave slots its registers allow (FP32 16 or 12, FP64 8 per CU).
template <typename R, int SH>
struct RowsCfg {
  static constexpr bool F32 = sizeof(R) == 4;
  static constexpr int TY = F32 ? (SH == 1 ? 6 : (SH == 2 ? 5 : (SH == 4 ? 6 : 4))) : (SH == 3 ? 4 : 2);
  static constexpr int ZS = SH == 1 ? 160 : (SH == 3 || SH == 4 ? 128 : 256);
  static constexpr int NTMAX = F32 ? (SH == 4 ? 768 : 1024) : 512;
  static constexpr int WPE = F32 ? (SH == 4 ? 3 : 4) : 2;
  static constexpr int NHT0 = F32 ? 1 : 2;  // halo entries per thread: NHT0 or NHT0 + 1
};

template <typename R>
static int rows_ty(int shape) {
  return shape == 1 ? RowsCfg<R, 1>::TY
                    : (shape == 2 ? RowsCfg<R, 2>::TY
                                  : (shape == 3 ? RowsCfg<R, 3>::TY : (shape == 4 ? RowsCfg<R, 4>::TY : RowsCfg<R, 0>::TY)));
}
template <typename R>
static int rows_zs(int shape) {
  return shape == 1 ? RowsCfg<R, 1>::ZS
                    : (shape == 2 ? RowsCfg<R, 2>::ZS
                                  : (shape == 3 ? RowsCfg<R, 3>::ZS : (shape == 4 ? RowsCfg<R, 4>::ZS : RowsCfg<R, 0>::ZS)));
}
template <typename R>
static int rows_ntmax(int shape) {
  return shape == 4 ? RowsCfg<R, 4>::NTMAX : RowsCfg<R, 0>::NTMAX;
}

template <typename R, int S, int NHT, int SH, bool WIDE = false>
static const void* rows_kernel_sh(int var) {
  using Cf = RowsCfg<R, SH>;
  constexpr int NB2 = sizeof(R) == 4 ? 2 : 1;  // an FP64 double-buffered tile exceeds the LDS
  if (var & 1) return (const void*)bulk_rows3d<R, R, S, Cf::TY, Cf::NTMAX, NHT, Cf::WPE, NB2, Cf::ZS, WIDE>;
  return (const void*)bulk_rows3d<R, R, S, Cf::TY, Cf::NTMAX, NHT, Cf::WPE, 1, Cf::ZS, WIDE>;
}

template <typename R, int S, int NHT>
static const void* rows_kernel(int var, int shape, bool wide) {
  if (wide) {  // one descriptor per field: 4-row workgroups (FP32 takes no row tiles at sbporder 4)
    if constexpr (sizeof(R) == 4 && S == 4) throw Error{"row tiles: no FP32 kernel at sbporder 4"};
    else return rows_kernel_sh<R, S, NHT, 0, true>(var);
  }
  if constexpr (sizeof(R) == 4) {
    if constexpr (S == 4) {
      if (shape == 4) return rows_kernel_sh<R, S, NHT, 4>(var);
    } else {
      if (shape == 1) return rows_kernel_sh<R, S, NHT, 1>(var);
      if (shape == 2) return rows_kernel_sh<R, S, NHT, 2>(var);
    }
  } else {
    if (shape == 3) return rows_kernel_sh<R, S, NHT, 3>(var);
  }
  return rows_kernel_sh<R, S, NHT, 0>(var);
}

template <typename R>
static const void* rows_kernel_for(int s, int nht, int var, int shape, bool wide) {
  constexpr int N0 = RowsCfg<R, 0>::NHT0;
  if (s == 2) return nht == N0 ? rows_kernel<R, 2, N0>(var, shape, wide) : rows_kernel<R, 2, N0 + 1>(var, shape, wide);
  if (s == 3) return nht == N0 ? rows_kernel<R, 3, N0>(var, shape, wide) : rows_kernel<R, 3, N0 + 1>(var, shape, wide);
  return nht == N0 ? rows_kernel<R, 4, N0>(var, shape, wide) : rows_kernel<R, 4, N0 + 1>(var, shape, wide);
}

// dynamic LDS of the row-tile kernel: [NBUF][8][TY + 2r][ZS]
template <typename R>
static size_t rows_lds(int s, int var, int shape) {
  const int nb = (sizeof(R) == 4 && (var & 1)) ? 2 : 1;
  return (size_t)nb * 8 * (rows_ty<R>(shape) + 2 * (s - 1)) * rows_zs<R>(shape) * sizeof(R);
}

template <typename R>
static void launch_rows(const Tiled3D& ta, int s, const R* f, R* fn, R* df, const double* hs, hipStream_t st) {
  const int nz = ta.hi[2] - ta.lo[2];
  if (ta.grp_y < 1 || ta.grp_x < 1) throw Error{"row tiles: patch shape not set"};
  const int nt = ((rows_ty<R>(ta.rows_shape) * nz + 63) / 64) * 64;
  const size_t lds = rows_lds<R>(s, ta.rows_var, ta.rows_shape);
  const dim3 grid((unsigned)(((ta.ntiles + 7) / 8) * 8));
  void* args[] = {(void*)&ta, (void*)&f, (void*)&fn, (void*)&df, (void*)&hs};
  HIPCHK(hipLaunchKernel(rows_kernel_for<R>(s, ta.rows_nht, ta.rows_var, ta.rows_shape, ta.wide != 0), grid, dim3(nt),
                         args, lds, st));
}

// The external SAT of a face normal to axis d, as the 9x9 map M[row][col] of
// a point's fields, only couples (v_d, s_dd) into v_d, s_dd, s_aa, s_bb and
// (v_c, s_dc) into v_c, s_dc for the two other axes c.  Extract those 16
// entries; false if any other entry is not negligible (non-Cartesian face).
static bool compact_face_sat(const double* M, int d, double* fc) {
  static const int sid[3][3] = {{3, 4, 5}, {4, 6, 7}, {5, 7, 8}};
  const int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
  bool used[81] = {};
  int n = 0;
  auto take = [&](int row, int col) {
    used[row * 9 + col] = true;
    fc[n++] = M[row * 9 + col];
  };
  for (int row : {d, sid[d][d], sid[a][a], sid[b][b]}) {
    take(row, d);
    take(row, sid[d][d]);
  }
  for (int c : {a, b})
    for (int row : {c, sid[d][c]}) {
      take(row, c);
      take(row, sid[d][c]);
    }
  double mx = 0.;
  for (int q = 0; q < 81; q++) mx = fmax(mx, fabs(M[q]));
  for (int q = 0; q < 81; q++)
    if (!used[q] && fabs(M[q]) > 1e-9 * mx) return false;
  return true;
}

// The same for a 2-D mode II face normal to axis d (fields vx vy sxx sxy
// syy): (v_d, s_dd) couple into v_d, s_dd, s_aa; (v_a, sxy) into v_a, sxy.
static bool compact_face_sat2(const double* M, int d, double* fc) {
  const int a = 1 - d, sdd = d == 0 ? 2 : 4, saa = d == 0 ? 4 : 2;
  bool used[25] = {};
  int n = 0;
  auto take = [&](int row, int col) {
    used[row * 5 + col] = true;
    fc[n++] = M[row * 5 + col];
  };
  for (int row : {d, sdd, saa}) {
    take(row, d);
    take(row, sdd);
  }
  for (int row : {a, 3}) {
    take(row, a);
    take(row, 3);
  }
  double mx = 0.;
  for (int q = 0; q < 25; q++) mx = fmax(mx, fabs(M[q]));
  for (int q = 0; q < 25; q++)
    if (!used[q] && fabs(M[q]) > 1e-9 * mx) return false;
  return true;
}

template <typename R>
Solver<R>::Solver(const std::string& path, const fdf_opts& o) {
  opts = o;
  HIPCHK(hipSetDevice(o.device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  // grid, metric and Jacobian built on the device (FDF_HOST_SETUP=1: on the
  // host, then uploaded)
  const char* eh = getenv("FDF_HOST_SETUP");
  if (!(eh && atoi(eh) != 0)) dgrid.reset(new DeviceGrid(stream));
  compile_problem(path, S, o.write_outputs != 0, o.quiet != 0, dgrid.get());
  // measured best tile heights on a 401x201x201 block: 4 rows (FP64), 8 (FP32)
  const int ty_default = sizeof(R) == 8 ? 4 : 8;
  tile_ty = ty_default;
  if (const char* e = getenv("FDF_TILE_TY")) tile_ty = atoi(e);
  if (tile_ty != 2 && tile_ty != 4 && tile_ty != 8) tile_ty = ty_default;
  if (const char* e = getenv("FDF_TILE2_TW")) tile2_tw = atoi(e), tile2_env = true;
  if (tile2_tw != 64 && tile2_tw != 128 && tile2_tw != 256) tile2_tw = 64;
  if (const char* e = getenv("FDF_TILE2_CX")) tile2_cx = atoi(e) > 0 ? atoi(e) : 2, tile2_env = true;
  for (int e = 0; e < 2 * NEV; e++) HIPCHK(hipEventCreate(&evs[e]));
  const size_t N = S.N;
  G.ndim = S.ndim;
  G.mode = S.mode;
  G.nf = S.nf;
  G.s = S.sbporder;
  G.Ny = S.nx[1];
  G.Nz = S.nx[2];
  G.N = (long long)N;
  G.plastic = S.plastic;
  G.hetmat = S.hetmat;
  G.hetstress = S.hetstress;
  G.cdiss = S.cdiss;
  for (int i = 0; i < 6; i++) G.s0[i] = S.s0[i];
  G.ns = S.ns;
  G.nv = S.nv;
  for (int i = 0; i < 6; i++) G.sidx[i] = S.sidx[i];
  G.nfp = S.nfp;
  G.ptensor = S.plastic_tensor;
  // SBP tables
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_fd), S.fd.c, sizeof(S.fd.c)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_diss), S.fd.d, sizeof(S.fd.d)));
  // grid arrays in the bulk precision: FP64 arrays are copied as they are,
  // FP32 ones converted on the device (round to nearest, as a host cast)
  auto upload_bulk = [&](DBuf<R>& dst, const double* h, size_t n) {
    if constexpr (std::is_same<R, double>::value) {
      dst.upload(h, n);
    } else {
      DBuf<double> tmp;
      tmp.upload(h, n);
      dst.alloc(n);
      if (n) {
        hipLaunchKernelGGL(convert_from_double<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, tmp.p,
                           dst.p, (long long)n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
      }
    }
  };
  if (dgrid) {
    if constexpr (std::is_same<R, double>::value) {
      metric = std::move(dgrid->metric);
      jac = std::move(dgrid->jac);
      metric64 = metric.p;
      jac64 = jac.p;
    } else {
      auto conv = [&](DBuf<R>& dst, const DBuf<double>& src) {
        dst.alloc(src.n);
        if (src.n) {
          hipLaunchKernelGGL(convert_from_double<R>, dim3((unsigned)((src.n + 255) / 256)), dim3(256), 0, stream, src.p,
                             dst.p, (long long)src.n);
          HIPCHK(hipGetLastError());
        }
      };
      conv(metric, dgrid->metric);
      conv(jac, dgrid->jac);
      HIPCHK(hipStreamSynchronize(stream));
      metric64 = dgrid->metric.p;
      jac64 = dgrid->jac.p;
    }
    x = std::move(dgrid->x);
  } else {
    upload_bulk(metric, S.metric.data(), S.metric.size());
    upload_bulk(jac, S.jac.data(), S.jac.size());
    x.upload(S.x.data(), S.x.size());
  }
  if (S.hetmat) upload_bulk(mat, S.mat.data(), S.mat.size());
  if (S.hetstress) hs.upload(S.s);
  for (int b = 0; b < 2; b++) {
    f[b].alloc((size_t)S.nf * N);
    f[b].zero(stream);
  }
  df.alloc((size_t)S.nf * N);
  df.zero(stream);
  if (S.nfp) {
    fp.alloc((size_t)S.nfp * N);
    fp.zero(stream);
  }
  err.alloc(1);
  err.zero(stream);
  dstep.alloc(1);
  dstep.zero(stream);
  // blocks
  const int nd = S.ndim;
  fbufs.resize(S.blocks.size() * 6);
  for (size_t bi = 0; bi < S.blocks.size(); bi++) {
    const BlockH& B = S.blocks[bi];
    BlockDev D{};
    for (int d = 0; d < 3; d++) {
      D.lo[d] = B.mlb[d];
      D.n[d] = B.nx[d];
      D.hi[d] = B.prb[d];
      D.mc[d] = B.mc[d];
      D.mrb[d] = B.mrb[d];
      D.dx[d] = B.dx[d];
      for (int q = 0; q < 4; q++) D.K[d][q] = B.K[d][q];
    }
    D.nu = 0.5 * B.lam / (B.lam + B.G);
    D.lam = B.lam;
    D.G = B.G;
    D.pmu = B.pmu;
    D.pc = B.pc;
    D.pbeta = B.pbeta;
    D.peta = B.peta;
    for (int fc = 0; fc < 2 * nd; fc++) {
      const FaceBC& F = B.face[fc];
      FaceDev& FD = D.face[fc];
      FD.active = F.active;
      if (!F.active) continue;
      FD.n0 = F.n0;
      FD.n1 = F.n1;
      FD.r = F.r;
      FD.alpha = F.alpha;
      FD.theta = F.theta;
      FD.cp = F.cp;
      FD.cs = F.cs;
      FD.zp = F.zp;
      FD.zs = F.zs;
      FD.gamma = F.gamma;
      FaceBufs& fb = fbufs[bi * 6 + fc];
      fb.nrm.upload(F.nrm);
      fb.dl.upload(F.dl);
      FD.nrm = fb.nrm.p;
      FD.dl = fb.dl.p;
    }
    bdev.push_back(D);
    // rectilinear fast paths: 3-D, and 2-D mode II without plasticity;
    // homogeneous material, no dissipation, not the bit-faithful FP64 mode.
    // The tiled kernels address the field arrays through buffer descriptors
    // with 32-bit offsets: one descriptor over all fields while nf N
    // elements fit in 4 GiB, else one per field (each field within 4 GiB;
    // FDF_WIDE=1 takes the per-field form at any size)
    const double fbytes = (double)N * sizeof(R);
    const bool fits = fbytes < 4294967296.;
    const char* ew = getenv("FDF_WIDE");
    const bool wide = (double)S.nf * fbytes >= 4294967296. || (ew && atoi(ew) != 0);
    const bool ok3 = nd == 3, ok2 = nd == 2 && S.mode == 2 && !S.plastic;
    const bool kind = !S.hetmat && !(S.cdiss > 0.) && !(sizeof(R) == 8 && o.exact_fp64);
    bool ok = (ok3 || ok2) && kind && fits;
    for (int d = 0; d < nd && ok; d++) ok = B.mrb[d] > B.mc[d] && B.nx[d] >= 3 * (S.sbporder - 1);
    // any other 2-D mode II block (curvilinear, or plastic) may take the
    // curvilinear tiled kernel, which reads the metric per point
    bool okc = nd == 2 && S.mode == 2 && kind && fits && !(getenv("FDF_CURV") && atoi(getenv("FDF_CURV")) == 0);
    for (int d = 0; d < nd && okc; d++) okc = B.mrb[d] > B.mc[d] && B.nx[d] >= 3 * (S.sbporder - 1);
    if (!fits && (ok3 || (nd == 2 && S.mode == 2)) && kind && !o.quiet)
      fprintf(stderr, "fdfault_amd: block %zu runs on the generic bulk kernel: one field of %lld points is "
              "above 4 GiB\n", bi, (long long)N);
    double mdiag[3] = {0., 0., 0.};
    if (ok) ok = dgrid ? dgrid->uniform(S, B, mdiag, metric64, jac64) : block_metric_uniform(S, B, mdiag);
    Tiled3D ta{};
    Tiled2D t2{};
    if (ok && ok3) {
      for (int d = 0; d < 3; d++) {
        ta.lo[d] = B.mlb[d];
        ta.hi[d] = B.prb[d];
        ta.mc[d] = B.mc[d];
        ta.mrb[d] = B.mrb[d];
      }
      ta.Ny = S.nx[1];
      ta.Nz = S.nx[2];
      ta.N = (long long)N;
      const int cx_default = 4;  // x planes per tile (short chunks: more workgroups in flight)
      ta.cx = cx_default;
      if (const char* e = getenv("FDF_TILE_CX")) ta.cx = atoi(e) > 0 ? atoi(e) : cx_default;
      ta.nzt = (ta.hi[2] - ta.lo[2] + 63) / 64;
      ta.nyt = (ta.hi[1] - ta.lo[1] + 3) / 4;
      ta.nxc = (ta.hi[0] - ta.lo[0] + ta.cx - 1) / ta.cx;
      ta.ntiles = ta.nzt * ta.nyt * ta.nxc;
      for (int d = 0; d < 3; d++) {
        ta.cv[d] = B.K[d][0] * mdiag[d];
        ta.cg2[d] = B.K[d][1] * mdiag[d];
        ta.cg[d] = B.K[d][2] * mdiag[d];
        ta.cl[d] = B.K[d][3] * mdiag[d];
        ta.fcv[d] = (float)ta.cv[d];
        ta.fcg2[d] = (float)ta.cg2[d];
        ta.fcg[d] = (float)ta.cg[d];
        ta.fcl[d] = (float)ta.cl[d];
      }
      ta.hetstress = S.hetstress;
      ta.wide = wide;
      for (int q = 0; q < 6; q++) ta.s0[q] = S.s0[q];
      for (int e = 0; e < 6; e++) ta.face_active[e] = D.face[e].active;
      // row-shaped tiles when TY whole rows fit one workgroup and its halo
      // needs at most NHT0 + 1 entries per thread (FDF_ROWS=0 keeps 64-wide
      // tiles; FDF_ROWS=1 single-buffered, 2 double-buffered LDS tile;
      // FDF_ROWS_SHAPE=0 keeps the 4-row FP32 and 2-row FP64 workgroups
      // for short rows)
      {
        const int nz = ta.hi[2] - ta.lo[2], r = S.sbporder - 1;
        const char* es = getenv("FDF_ROWS_SHAPE");
        int shape = 0;  // (per-field descriptors: 4-row workgroups only)
        if (sizeof(R) == 4 && !(es && atoi(es) == 0) && !wide)
          shape = nz <= RowsCfg<R, 1>::ZS ? 1 : (RowsCfg<R, 2>::TY * nz <= 1024 ? 2 : 0);
        if (sizeof(R) == 8 && !(es && atoi(es) == 0) && !wide && nz <= RowsCfg<R, 3>::ZS) shape = 3;
        // FP32 sbporder 4 takes row tiles only in the 3-wave shape (rows of <= 128)
        const bool f32s4 = sizeof(R) == 4 && S.sbporder == 4;
        if (f32s4) shape = (!wide && nz <= RowsCfg<R, 4>::ZS && RowsCfg<R, 4>::TY * nz <= RowsCfg<R, 4>::NTMAX) ? 4 : -1;
        const int TYs = rows_ty<R>(shape), NTM = rows_ntmax<R>(shape);
        const int N0 = RowsCfg<R, 0>::NHT0;
        const int nt = ((TYs * nz + 63) / 64) * 64;
        const int nht = (2 * r * nz + nt - 1) / nt;
        int dev = 0, maxlds = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
        const char* ev = getenv("FDF_ROWS");
        const bool want = !(ev && atoi(ev) == 0);
        int var = rows_var_default;
        if (ev && atoi(ev) > 0) var = (atoi(ev) - 1) & 1;
        if (rows_lds<R>(S.sbporder, var, shape) + 1024 > (size_t)maxlds) var &= ~1;
        const size_t lds = rows_lds<R>(S.sbporder, var, shape);
        // (FP32 with sbporder 4 exceeds the 128-register budget of 1024 threads)
        const bool fits = shape >= 0;
        if (want && fits && nz <= 256 && nt <= NTM && nht <= N0 + 1 && lds + 1024 <= (size_t)maxlds) {
          ta.rows_nht = std::max(nht, N0);
          ta.rows_var = var;
          ta.rows_shape = shape;
          ta.nyt = (ta.hi[1] - ta.lo[1] + TYs - 1) / TYs;
          HIPCHK(hipFuncSetAttribute(rows_kernel_for<R>(S.sbporder, ta.rows_nht, var, shape, wide),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
          ta.ntiles = ta.nyt * ta.nxc;
          // tile patches of grp_y x grp_x: one XCD's worth of workgroups
          // (FDF_ROWS_GROUP="gy,gx"; "1000,1" is the y-fastest order)
          ta.grp_y = rows_grp_y;
          ta.grp_x = rows_grp_x;
          if (const char* eg = getenv("FDF_ROWS_GROUP")) {
            int gy = 0, gx = 0;
            if (sscanf(eg, "%d,%d", &gy, &gx) == 2 && gy > 0 && gx > 0) {
              ta.grp_y = gy;
              ta.grp_x = gx;
            }
          }
        }
      }
    }
    if ((ok && ok2) || okc) {
      for (int d = 0; d < 2; d++) {
        t2.lo[d] = B.mlb[d];
        t2.hi[d] = B.prb[d];
        t2.mc[d] = B.mc[d];
        t2.mrb[d] = B.mrb[d];
      }
      t2.Ny = S.nx[1];
      t2.N = (long long)N;
      t2.cx = tile2_cx;
      t2.tw = wide ? 64 : tile2_tw;  // (per-field descriptors: 64-wide tiles only)
      t2.wide = wide;
      if (!tile2_env) {
        // 64-wide tiles, 2-plane chunks (short marches keep the DRAM access
        // stream dense: measured best on 2x1601x801 and 2x801x401); blocks
        // that would give fewer than 16 workgroups per CU march one plane
        int dev = 0, ncu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const long long nx = t2.hi[0] - t2.lo[0], ny = t2.hi[1] - t2.lo[1];
        if (((ny + t2.tw - 1) / t2.tw) * ((nx + t2.cx - 1) / t2.cx) < 16LL * ncu) t2.cx = 1;
      }
      t2.nyt = (t2.hi[1] - t2.lo[1] + t2.tw - 1) / t2.tw;
      t2.nxc = (t2.hi[0] - t2.lo[0] + t2.cx - 1) / t2.cx;
      t2.ntiles = t2.nyt * t2.nxc;
      for (int d = 0; d < 2; d++) {
        t2.cv[d] = B.K[d][0] * mdiag[d];
        t2.cg2[d] = B.K[d][1] * mdiag[d];
        t2.cg[d] = B.K[d][2] * mdiag[d];
        t2.cl[d] = B.K[d][3] * mdiag[d];
        t2.fcv[d] = (float)t2.cv[d];
        t2.fcg2[d] = (float)t2.cg2[d];
        t2.fcg[d] = (float)t2.cg[d];
        t2.fcl[d] = (float)t2.cl[d];
      }
      t2.hetstress = S.hetstress;
      t2.s0[0] = S.s0[0];
      t2.s0[1] = S.s0[1];
      t2.s0[2] = S.s0[3];
      for (int e = 0; e < 4; e++) t2.face_active[e] = D.face[e].active;
    }
    tiled_ok.push_back(ok);
    curv_ok.push_back(okc);
    tiles.push_back(ta);
    tiles2.push_back(t2);
  }
  // interfaces
  ibufs.resize(S.ifaces.size());
  for (size_t ii = 0; ii < S.ifaces.size(); ii++) {
    const IfaceH& I = S.ifaces[ii];
    IfBufs& ib = ibufs[ii];
    FaultDev F{};
    F.law = I.law;
    F.dir = I.dir;
    F.a = (I.dir == 0) ? 1 : 0;
    F.b = (I.dir == 2) ? 1 : 2;
    F.n0 = I.n0;
    F.n1 = I.n1;
    F.n = I.n0 * I.n1;
    for (int d = 0; d < 3; d++) F.org[d] = I.org[d];
    F.cp1 = I.cp1; F.cs1 = I.cs1; F.zp1 = I.zp1; F.zs1 = I.zs1;
    F.cp2 = I.cp2; F.cs2 = I.cs2; F.zp2 = I.zp2; F.zs2 = I.zs2;
    F.gamma1 = I.gamma1;
    F.gamma2 = I.gamma2;
    ib.nrm.upload(I.nrm);
    ib.dl1.upload(I.dl1);
    ib.dl2.upload(I.dl2);
    F.nrm = ib.nrm.p;
    F.dl1 = ib.dl1.p;
    F.dl2 = ib.dl2.p;
    const size_t n = (size_t)F.n;
    auto pack = [](const std::vector<PertH>& P, DBuf<double>& xy, DBuf<double>& par) {
      std::vector<double> hx, hp;
      for (const PertH& p : P) {
        hx.insert(hx.end(), p.xy.begin(), p.xy.end());
        hp.push_back(p.t0);
        for (int q = 0; q < 9; q++) hp.push_back(p.amp[q]);
      }
      xy.upload(hx);
      par.upload(hp);
    };
    F.nloads = (int)I.loads.size();
    pack(I.loads, ib.load_xy, ib.load_par);
    F.load_xy = ib.load_xy.p;
    F.load_par = ib.load_par.p;
    F.load_file = I.load_file;
    if (I.load_file) {
      std::vector<double> all;
      for (int q = 0; q < 3; q++) all.insert(all.end(), I.lf[q].begin(), I.lf[q].end());
      ib.lf.upload(all);
      F.lf = ib.lf.p;
    }
    F.nperts = (int)I.perts.size();
    pack(I.perts, ib.pert_xy, ib.pert_par);
    F.pert_xy = ib.pert_xy.p;
    F.pert_par = ib.pert_par.p;
    F.param_file = I.param_file;
    if (I.param_file) {
      std::vector<double> all;
      for (int q = 0; q < 9; q++) all.insert(all.end(), I.pf[q].begin(), I.pf[q].end());
      ib.pf.upload(all);
      F.pf = ib.pf.p;
    }
    const size_t nc = (size_t)(S.ndim - 1) * n;
    for (DBuf<double>* b : {&ib.u, &ib.du, &ib.v, &ib.s, &ib.sn}) { b->alloc(n); b->zero(stream); }
    for (DBuf<double>* b : {&ib.ux, &ib.dux, &ib.vx, &ib.sx}) { b->alloc(nc); b->zero(stream); }
    ib.state.alloc(n);
    ib.state.zero(stream);
    ib.dstate.alloc(n);
    ib.dstate.zero(stream);
    if (I.law == LAW_STZ) HIPCHK(hipMemcpy(ib.state.p, I.state0.data(), n * sizeof(double), hipMemcpyHostToDevice));
    ib.sat.alloc((size_t)4 * S.nf * n);
    ib.sat.zero(stream);
    F.u = ib.u.p; F.du = ib.du.p; F.v = ib.v.p; F.s = ib.s.p; F.sn = ib.sn.p;
    F.ux = ib.ux.p; F.dux = ib.dux.p; F.vx = ib.vx.p; F.sx = ib.sx.p;
    F.state = ib.state.p; F.dstate = ib.dstate.p;
    F.sat = ib.sat.p;
    if (S.has_front && I.law != LAW_LOCKED) {
      std::vector<double> m1(n, -1.);
      ib.front.upload(m1);
      F.front = ib.front.p;
      F.front_value = S.front_value;
      F.front_field = S.front_field;
    }
    if (I.law == LAW_STZ) {
      // FDF_STZ_NEWTON=exact: plain solves with the exact derivative instead
      // of the reference's (to separate the two effects in comparisons)
      if (const char* e = getenv("FDF_STZ_NEWTON")) F.stz_exact_newton = !strcmp(e, "exact");
      if (!stz_cnt.p) {
        stz_cnt.alloc(5);
        stz_cnt.zero(stream);
      }
      F.stats = stz_cnt.p + 1;
    }
    fdev.push_back(F);
    // register this interface on the two blocks' face lists (interface order)
    for (int side = 0; side < 2; side++) {
      BlockDev& D = bdev[side == 0 ? I.b1 : I.b2];
      if (D.nif >= 6) throw Error{"too many interfaces on one block"};
      IfFaceDev& X = D.ifc[D.nif++];
      X.side = side;
      X.dir = I.dir;
      X.fixed = I.org[I.dir] + side;
      X.a = F.a;
      X.b = F.b;
      X.org_a = I.org[F.a];
      X.org_b = I.org[F.b];
      X.n = F.n;
      X.n1 = F.n1;
      X.sat = ib.sat.p;
    }
  }
  // rectilinear blocks: interface faces and external SAT maps for the tiled kernels
  {
    const int nfu = S.nf;
    std::vector<double> unit((size_t)nfu * nfu, 0.);
    for (int q = 0; q < nfu; q++) unit[q * nfu + q] = 1.;  // field q of point q
    DBuf<double> dunit, fm;
    dunit.upload(unit);
    fm.alloc(81);
    fcoefs.resize(bdev.size());
    for (size_t b = 0; b < bdev.size(); b++) {
      if (!tiled_ok[b]) continue;
      const BlockDev& D = bdev[b];
      const int nfc = 2 * nd, ne = nd == 3 ? 16 : 10;  // faces, compact entries per face
      double fc[6][16] = {};
      for (int e = 0; e < nfc && tiled_ok[b]; e++) {
        if (!D.face[e].active) continue;
        if (nd == 3)
          hipLaunchKernelGGL(face_matrix_kernel<2>, dim3(1), dim3(64), 0, stream, D.face[e], G, S.dt,
                             D.face[e].n0 / 2, D.face[e].n1 / 2, dunit.p, fm.p);
        else
          hipLaunchKernelGGL(face_matrix_kernel<0>, dim3(1), dim3(64), 0, stream, D.face[e], G, S.dt,
                             D.face[e].n0 / 2, D.face[e].n1 / 2, dunit.p, fm.p);
        HIPCHK(hipGetLastError());
        double M[81];
        HIPCHK(hipMemcpyAsync(M, fm.p, sizeof(double) * nfu * nfu, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        const bool good = nd == 3 ? compact_face_sat(M, e >> 1, fc[e]) : compact_face_sat2(M, e >> 1, fc[e]);
        if (!good) tiled_ok[b] = false;
      }
      if (!tiled_ok[b]) continue;
      std::vector<R> fcr((size_t)nfc * ne);
      for (int q = 0; q < nfc * ne; q++) fcr[q] = (R)fc[q / ne][q % ne];
      fcoefs[b].upload(fcr);
      if (nd == 3) {
        Tiled3D& ta = tiles[b];
        ta.fcoef = fcoefs[b].p;
        ta.nif = D.nif;
        for (int e = 0; e < D.nif; e++) {
          const IfFaceDev& X = D.ifc[e];
          ta.if_dir[e] = X.dir;
          ta.if_fixed[e] = X.fixed;
          ta.if_a[e] = X.a;
          ta.if_b[e] = X.b;
          ta.if_orga[e] = X.org_a;
          ta.if_orgb[e] = X.org_b;
          ta.if_n[e] = X.n;
          ta.if_n1[e] = X.n1;
          ta.if_side[e] = X.side;
          ta.if_sat[e] = X.sat;
        }
      } else {
        Tiled2D& t2 = tiles2[b];
        t2.fcoef = fcoefs[b].p;
        if (D.nif > 4) throw Error{"too many interfaces on one block"};
        t2.nif = D.nif;
        for (int e = 0; e < D.nif; e++) {
          const IfFaceDev& X = D.ifc[e];
          t2.if_dir[e] = X.dir;
          t2.if_fixed[e] = X.fixed;
          t2.if_a[e] = X.a;
          t2.if_orga[e] = X.org_a;
          t2.if_n[e] = X.n;
          t2.if_side[e] = X.side;
          t2.if_sat[e] = X.sat;
        }
      }
    }
  }
  HIPCHK(hipStreamSynchronize(stream));
  smaps.resize(curv_ok.size() * 4);  // (sized once: the maps' device pointers are kept)
  for (size_t b = 0; b < curv_ok.size(); b++) {
    // the rectilinear kernel is preferred where it applies
    curv_ok[b] = curv_ok[b] && !tiled_ok[b] && bdev[b].nif <= 4;
    if (!curv_ok[b]) continue;
    const BlockDev& D = bdev[b];
    Tiled2D& t2 = tiles2[b];
    t2.nif = D.nif;
    for (int e = 0; e < D.nif; e++) {
      const IfFaceDev& X = D.ifc[e];
      t2.if_dir[e] = X.dir;
      t2.if_fixed[e] = X.fixed;
      t2.if_a[e] = X.a;
      t2.if_orga[e] = X.org_a;
      t2.if_n[e] = X.n;
      t2.if_side[e] = X.side;
      t2.if_sat[e] = X.sat;
    }
    // the SAT of every external face point as a [rows][5] map (rows: the
    // evolved fields, szz included for plastic mode II)
    const int nr = S.nf;
    std::vector<double> unit(25, 0.);
    for (int q = 0; q < 5; q++) unit[q * 5 + q] = 1.;
    DBuf<double> dunit5;
    dunit5.upload(unit);
    for (int fc = 0; fc < 4; fc++) {
      t2.smap[fc] = nullptr;
      if (!D.face[fc].active) continue;
      if (D.face[fc].n1 != 1) throw Error{"2-D face with more than one point across"};
      const int n = D.face[fc].n0;
      DBuf<double> out;
      out.alloc((size_t)n * nr * 5);
      hipLaunchKernelGGL(face_map_kernel<0>, dim3((n * 5 + 255) / 256), dim3(256), 0, stream, D.face[fc], G, S.dt, n,
                         nr, dunit5.p, out.p);
      HIPCHK(hipGetLastError());
      std::vector<double> h((size_t)n * nr * 5);
      HIPCHK(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      std::vector<R> hr(h.size());
      for (size_t q = 0; q < h.size(); q++) hr[q] = (R)h[q];
      smaps[b * 4 + fc].upload(hr);
      t2.smap[fc] = smaps[b * 4 + fc].p;
    }
  }
  {
    bool allt = nd == 2 && S.blocks.size() > 1, allc = allt;
    for (size_t b = 0; b < S.blocks.size() && (allt || allc); b++) {
      allt = allt && tiled_ok[b] && tiles2[b].tw == tiles2[0].tw;
      allc = allc && curv_ok[b] && tiles2[b].tw == tiles2[0].tw;
    }
    const char* em = getenv("FDF_MERGE2D");
    if ((allt || allc) && !(em && atoi(em) == 0)) {
      merged = allt ? 1 : 2;
      std::vector<int2> wl;
      for (size_t b = 0; b < S.blocks.size(); b++)
        for (int tl = 0; tl < tiles2[b].ntiles; tl++) wl.push_back(make_int2((int)b, tl));
      mwork.upload(wl);
      mnwork = (int)wl.size();
      mdesc.resize((size_t)S.rk.nstages * 2);
      for (int st = 0; st < S.rk.nstages; st++)
        for (int rt = 0; rt < 2; rt++) {
          const StageDev sd{S.rk.A[st], S.rk.B[st], S.dt, rt};
          std::vector<Tiled2D> d;
          for (size_t b = 0; b < S.blocks.size(); b++) d.push_back(stage_tile2((int)b, sd));
          mdesc[(size_t)st * 2 + rt].upload(d);
        }
      if (merged == 2) mbd.upload(bdev);
    }
    // on by default (FDF_FUSE2D=0 turns it off): the fault code (its lean
    // form, about 135 VGPRs) sets the fused kernel's registers, 3 waves per
    // SIMD, which the tiles of a single problem lose less to than a separate
    // fault launch costs
    const char* ef = getenv("FDF_FUSE2D");
    const bool want = ef ? atoi(ef) != 0 : true;
    if (merged == 1 && !tiles2[0].wide && S.ifaces.size() == 1 && want) {
      const IfaceH& I = S.ifaces[0];
      std::vector<FaultPair2D> pairs;
      int yA = 0, yB = 0;
      if (fault_chunks_2d(I, fdev[0], tiles2[I.b1], tiles2[I.b2], tiles2[0].tw, 0, I.b1, I.b2, pairs, yA, yB)) {
        std::vector<int2> tl;
        for (size_t b = 0; b < S.blocks.size(); b++)
          for (int t = 0; t < tiles2[b].ntiles; t++) {
            const int yt = t % tiles2[b].nyt;
            if (((int)b == I.b1 && yt == yA) || ((int)b == I.b2 && yt == yB)) continue;
            tl.push_back(make_int2((int)b, t));
          }
        int G0 = 0;
        const std::vector<int2> wlf = fused_work_list(tl, (int)pairs.size(), G0);
        mnfwork = (int)wlf.size();
        mfwork.upload(wlf);
        FaultBatchDev fb{};
        fb.F = fdev[0];
        fb.g = G;
        fb.f[0] = f[0].p;
        fb.f[1] = f[1].p;
        fb.mat = mat.p;
        fb.hs = hs.p;
        fb.dt = S.dt;
        mfdev.upload(std::vector<FaultBatchDev>{fb});
        mpairs.upload(pairs);
        fused2d = true;
        nfused = 1;
      }
    }
  }
  {
    // by default only for blocks of fewer than 8 rounds of one workgroup per
    // CU, where a launch's tail is a sizeable part of it (FDF_PAIR3D: 0 off,
    // 1 always)
    const char* ep = getenv("FDF_PAIR3D");
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const bool two = nd == 3 && S.blocks.size() == 2 && tiled_ok[0] && tiled_ok[1];
    const bool small = two && tiles[0].ntiles < 8 * ncu && tiles[1].ntiles < 8 * ncu;
    if (two && (ep ? atoi(ep) != 0 : small)) {
      pair3 = true;
      HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
  }
  for (size_t b = 0; b < tiled_ok.size(); b++) {
    ncurv += curv_ok[b] ? 1 : 0;
    ntiled += tiled_ok[b] ? 1 : 0;
    nwide += ((tiled_ok[b] && (nd == 3 ? tiles[b].wide : tiles2[b].wide)) || (curv_ok[b] && tiles2[b].wide)) ? 1 : 0;
    nrows += (tiled_ok[b] && nd == 3 && tiles[b].rows_nht > 0) ? 1 : 0;
  }
  // t = 0: interface solve without SAT (domain.cpp:141-143) on f = 0
  {
    FaultStageDev fs{0., 0., 0., 0., 1, -1., nullptr, 0., 0};
    for (size_t i = 0; i < S.ifaces.size(); i++)
      if (S.ifaces[i].law != LAW_LOCKED) launch_fault((int)i, fs);
  }
  // set_stress / remove_stress around the t = 0 output (problem.cpp:74-76):
  // f = ((0 + s0) + s - s0) - s, non-zero only with heterogeneous stress
  if (S.hetstress) {
    std::vector<double> h((size_t)S.nf * N, 0.);
    for (int i = 0; i < S.ns; i++) {
      const double s0 = S.s0[S.sidx[i]];
      double* fq = h.data() + (size_t)(S.nv + i) * N;
      const double* sq = S.s.data() + (size_t)i * N;
      for (size_t j = 0; j < N; j++) {
        double v = fq[j];
        v += s0;
        v += sq[j];
        v -= s0;
        v -= sq[j];
        fq[j] = v;
      }
    }
    HIPCHK(hipStreamSynchronize(stream));
    write_array(0, h.data(), (long)h.size());
  }
  // outputs and front at t = 0 (problem.cpp:66-82)
  if (opts.write_outputs) {
    const char endian = [] {
      const uint32_t v = 0x01020304;
      unsigned char c[4];
      memcpy(c, &v, 4);
      return c[0] == 4 ? '<' : (c[0] == 1 ? '>' : '=');
    }();
    for (const OutputH& U : S.outputs) {
      OutFile of;
      const std::string base = S.datadir + S.name + "_" + U.name;
      of.data = fopen((base + "_" + U.field + ".dat").c_str(), "wb");
      if (!of.data) throw Error{"Error opening file in outputunit.cpp"};
      static const char* xyz[3] = {"x", "y", "z"};
      for (int c = 0; c < S.ndim; c++) {
        FILE* xf = fopen((base + "_" + xyz[c] + ".dat").c_str(), "wb");
        if (!xf) throw Error{"Error opening file in outputunit.cpp"};
        std::vector<double> buf;
        coords_box(c, U.xm, U.xp, U.xs, buf);
        fwrite(buf.data(), sizeof(double), buf.size(), xf);
        fclose(xf);
      }
      if (U.tp > U.tm) {
        of.time = fopen((base + "_t.dat").c_str(), "wb");
        if (!of.time) throw Error{"Error opening file in outputunit.cpp"};
      }
      const int ntout = (U.tp - U.tm) / U.ts + 1;
      FILE* m = fopen((base + ".m").c_str(), "w");
      if (!m) throw Error{"Error writing parameters to matlab file"};
      fprintf(m, "endian = '%c';\nfield = '%s';\nnt = %d;\nnx = %d;\nny = %d;\nnz = %d;\n",
              endian == '<' ? 'l' : (endian == '>' ? 'b' : 'n'), U.field.c_str(), ntout, U.nx[0], U.nx[1], U.nx[2]);
      fclose(m);
      FILE* py = fopen((base + ".py").c_str(), "w");
      if (!py) throw Error{"Error writing parameters to python file"};
      fprintf(py, "endian = '%c'\nfield = '%s';\nnt = %d\nnx = %d\nny = %d\nnz = %d\n", endian, U.field.c_str(), ntout,
              U.nx[0], U.nx[1], U.nx[2]);
      fclose(py);
      outs.push_back(of);
    }
    setup_output_staging();
    write_outputs(0);
  }
  // front at t = 0 (front.cpp:103-125) from V of the initial interface solve
  if (S.has_front && S.front_field == 0) {
    for (size_t i = 0; i < S.ifaces.size(); i++) {
      if (S.ifaces[i].law == LAW_LOCKED) continue;
      IfBufs& ib = ibufs[i];
      std::vector<double> v(ib.v.n), tv(ib.v.n);
      HIPCHK(hipStreamSynchronize(stream));
      HIPCHK(hipMemcpy(v.data(), ib.v.p, v.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < v.size(); q++) tv[q] = (v[q] >= S.front_value) ? 0. : -1.;
      HIPCHK(hipMemcpy(ib.front.p, tv.data(), tv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  sync_check();
}

template <typename R>
Solver<R>::~Solver() {
  try {
    drain_outputs();
  } catch (...) {
  }
  for (OutFile& o : outs) {
    if (o.data) fclose(o.data);
    if (o.time) fclose(o.time);
  }
  if (stream) (void)hipStreamSynchronize(stream);
  if (ostream) (void)hipStreamSynchronize(ostream);
  for (OutStage& O : ostage) {
    for (hipEvent_t e : O.gathered) (void)hipEventDestroy(e);
    for (hipEvent_t e : O.copied) (void)hipEventDestroy(e);
    if (O.host) (void)hipHostFree(O.host);
  }
  if (ostream) (void)hipStreamDestroy(ostream);
  if (side) (void)hipStreamSynchronize(side);
  for (int e = 0; e < 2 * NEV; e++)
    if (evs[e]) (void)hipEventDestroy(evs[e]);
  for (int q = 0; q < 2; q++)
    if (graphs[q]) (void)hipGraphExecDestroy(graphs[q]);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

// the message of the fault kernels' error flags: the reference's text where
// it stops (stz.cpp:341-344, utilities.cpp:135-136,173-176)
static std::string fault_error_message(int e) {
  if (e & 16) return "Slip velocity exceeds melting value in stz.cpp";
  if (e & 1) return "zero derivative in Newton's method in utilities.cpp";
  if (e & 2) return "Newton's method failed to converge in utilities.cpp";
  if (e & 8) return "STZ effective temperature must be positive";
  return "bracket repair in Newton's method did not terminate";
}

template <typename R>
void Solver<R>::sync_check() {
  HIPCHK(hipStreamSynchronize(stream));
  int e = 0;
  HIPCHK(hipMemcpy(&e, err.p, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    int z = 0;
    HIPCHK(hipMemcpy(err.p, &z, sizeof(int), hipMemcpyHostToDevice));
    throw Error{fault_error_message(e)};
  }
}

template <typename R>
void Solver<R>::launch_fault(int i, const FaultStageDev& fs) {
  const FaultDev& F = fdev[i];
  const int nb = (F.n + 255) / 256;
  const R* fc = f[cur].p;
  const void* k = fault_kernel_for<R>(mode_id(), F.law);
  void* args[] = {(void*)&F, (void*)&G, (void*)&fs, (void*)&fc, (void*)&mat.p, (void*)&hs.p, (void*)&err.p, nullptr};
  int* no_iters = nullptr;
  args[7] = (void*)&no_iters;
  HIPCHK(hipLaunchKernel(k, dim3(nb), dim3(256), args, 0, stream));
  HIPCHK(hipGetLastError());
}

template <int M, typename R>
static void bulk_dispatch(bool diss, bool het, const BlockDev& D, const Boxes& bx, const GridDev& G, const StageDev& sd,
                          const R* f, R* fn, R* df, const R* mt, const R* jac, const R* mat, const double* hs,
                          hipStream_t s) {
  const long long np = bx.start[bx.nbox];
  if (np == 0) return;
  const dim3 grid((unsigned)((((np + 255) / 256) + 7) / 8 * 8)), blk(256);
  if (diss) {
    if (het) hipLaunchKernelGGL((bulk_kernel<M, R, R, true, true>), grid, blk, 0, s, D, bx, G, sd, f, fn, df, mt, jac, mat, hs);
    else hipLaunchKernelGGL((bulk_kernel<M, R, R, true, false>), grid, blk, 0, s, D, bx, G, sd, f, fn, df, mt, jac, mat, hs);
  } else {
    if (het) hipLaunchKernelGGL((bulk_kernel<M, R, R, false, true>), grid, blk, 0, s, D, bx, G, sd, f, fn, df, mt, jac, mat, hs);
    else hipLaunchKernelGGL((bulk_kernel<M, R, R, false, false>), grid, blk, 0, s, D, bx, G, sd, f, fn, df, mt, jac, mat, hs);
  }
}

// whole block as one box, or the six shell boxes around [mc, mrb)
static Boxes make_boxes(const BlockDev& D, bool shell) {
  Boxes b{};
  auto add = [&](int l0, int h0, int l1, int h1, int l2, int h2) {
    if (h0 <= l0 || h1 <= l1 || h2 <= l2) return;
    const int q = b.nbox++;
    b.lo[q][0] = l0; b.lo[q][1] = l1; b.lo[q][2] = l2;
    b.n[q][0] = h0 - l0; b.n[q][1] = h1 - l1; b.n[q][2] = h2 - l2;
  };
  if (!shell) {
    add(D.lo[0], D.hi[0], D.lo[1], D.hi[1], D.lo[2], D.hi[2]);
  } else {
    add(D.lo[0], D.mc[0], D.lo[1], D.hi[1], D.lo[2], D.hi[2]);
    add(D.mrb[0], D.hi[0], D.lo[1], D.hi[1], D.lo[2], D.hi[2]);
    add(D.mc[0], D.mrb[0], D.lo[1], D.mc[1], D.lo[2], D.hi[2]);
    add(D.mc[0], D.mrb[0], D.mrb[1], D.hi[1], D.lo[2], D.hi[2]);
    add(D.mc[0], D.mrb[0], D.mc[1], D.mrb[1], D.lo[2], D.mc[2]);
    add(D.mc[0], D.mrb[0], D.mc[1], D.mrb[1], D.mrb[2], D.hi[2]);
  }
  b.start[0] = 0;
  for (int q = 0; q < b.nbox; q++) b.start[q + 1] = b.start[q] + (long long)b.n[q][0] * b.n[q][1] * b.n[q][2];
  for (int q = b.nbox + 1; q < 7; q++) b.start[q] = b.start[b.nbox];
  return b;
}

// interior kernel launch; the compute type is FP64 for FP64 storage and
// FP32 for FP32 storage
template <typename R, int TY, bool WIDE = false>
static void launch_tiled_ty(const Tiled3D& ta, int s, const R* f, R* fn, R* df, const double* hs, hipStream_t st) {
  const dim3 grid((unsigned)(((ta.ntiles + 7) / 8) * 8)), blk(64 * TY);
  // waves per SIMD the register allocation must allow: 4 (<= 128 VGPRs)
  // for FP32; FP64 holds twice the registers per value and runs at 2
  constexpr int WPE = sizeof(R) == 4 ? 4 : 2;
  if (s == 2) hipLaunchKernelGGL((bulk_tiled3d<R, R, 2, TY, WPE, WIDE>), grid, blk, 0, st, ta, f, fn, df, hs);
  else if (s == 3) hipLaunchKernelGGL((bulk_tiled3d<R, R, 3, TY, WPE, WIDE>), grid, blk, 0, st, ta, f, fn, df, hs);
  else hipLaunchKernelGGL((bulk_tiled3d<R, R, 4, TY, WPE, WIDE>), grid, blk, 0, st, ta, f, fn, df, hs);
}

// (WIDE: one descriptor per field; instantiated for 64-wide tiles only)
template <typename R, int TW, bool WIDE = false>
static void launch_tiled2_tw(const Tiled2D& t2, int s, int par, hipStream_t st) {
  const dim3 grid((unsigned)(((t2.ntiles + 7) / 8) * 8)), blk(TW);
  if (s == 2) hipLaunchKernelGGL((bulk_tiled2d<R, R, 2, TW, WIDE>), grid, blk, 0, st, t2, par);
  else if (s == 3) hipLaunchKernelGGL((bulk_tiled2d<R, R, 3, TW, WIDE>), grid, blk, 0, st, t2, par);
  else hipLaunchKernelGGL((bulk_tiled2d<R, R, 4, TW, WIDE>), grid, blk, 0, st, t2, par);
}

template <typename R, int TW, bool PL, bool WIDE>
static void launch_curv2_pl(const Tiled2D& t2, const BlockDev& D, const R* mt, const R* jac, int s, int par,
                            hipStream_t st) {
  const dim3 grid((unsigned)(((t2.ntiles + 7) / 8) * 8)), blk(TW);
  if (s == 2) hipLaunchKernelGGL((bulk_curv2d<R, R, 2, TW, PL, WIDE>), grid, blk, 0, st, t2, D, mt, jac, par);
  else if (s == 3) hipLaunchKernelGGL((bulk_curv2d<R, R, 3, TW, PL, WIDE>), grid, blk, 0, st, t2, D, mt, jac, par);
  else hipLaunchKernelGGL((bulk_curv2d<R, R, 4, TW, PL, WIDE>), grid, blk, 0, st, t2, D, mt, jac, par);
}

template <typename R, int TW, bool WIDE = false>
static void launch_curv2_tw(const Tiled2D& t2, const BlockDev& D, const R* mt, const R* jac, int s, bool plastic,
                            int par, hipStream_t st) {
  if (plastic) launch_curv2_pl<R, TW, true, WIDE>(t2, D, mt, jac, s, par, st);
  else launch_curv2_pl<R, TW, false, WIDE>(t2, D, mt, jac, s, par, st);
}

// merged launches over every block of a 2-D problem, one workgroup per tile
template <typename R, int TW, bool WIDE = false>
static void launch_merged_tw(const Tiled2D* d, const int2* wk, int nw, int s, int par, int dfm, dim3 grid,
                             hipStream_t st) {
  if (s == 2) hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 2, TW, WIDE>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm);
  else if (s == 3)
    hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 3, TW, WIDE>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm);
  else hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 4, TW, WIDE>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm);
}

// stage launch of a 2-D problem with its fault solve (slip weakening)
template <typename R, int TW>
static void launch_fused_tw(const Tiled2D* d, const int2* wk, int nw, int s, int par, int dfm, dim3 grid,
                            const FaultBatchDev* fdv, const FaultPair2D* pairs, const FaultStageDev& fs, int* err,
                            hipStream_t st) {
  if (s == 2)
    hipLaunchKernelGGL((bulk_fault2d_batch<R, R, 2, TW, LAW_SLIPWEAK>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm,
                       fdv, pairs, fs, err);
  else if (s == 3)
    hipLaunchKernelGGL((bulk_fault2d_batch<R, R, 3, TW, LAW_SLIPWEAK>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm,
                       fdv, pairs, fs, err);
  else
    hipLaunchKernelGGL((bulk_fault2d_batch<R, R, 4, TW, LAW_SLIPWEAK>), grid, dim3(TW), 0, st, d, wk, nw, par, dfm,
                       fdv, pairs, fs, err);
}

template <typename R>
static void launch_fused(int tw, const Tiled2D* d, const int2* wk, int nw, int s, int par, int dfm, dim3 grid,
                         const FaultBatchDev* fdv, const FaultPair2D* pairs, const FaultStageDev& fs, int* err,
                         hipStream_t st) {
  if (tw == 64) launch_fused_tw<R, 64>(d, wk, nw, s, par, dfm, grid, fdv, pairs, fs, err, st);
  else if (tw == 128) launch_fused_tw<R, 128>(d, wk, nw, s, par, dfm, grid, fdv, pairs, fs, err, st);
  else launch_fused_tw<R, 256>(d, wk, nw, s, par, dfm, grid, fdv, pairs, fs, err, st);
}

template <typename R, int TW, bool PL, bool WIDE>
static void launch_mcurv_pl(const Tiled2D* d, const BlockDev* bds, const int2* wk, int nw, const R* mt, const R* jac,
                            int s, int par, int dfm, dim3 grid, hipStream_t st) {
  if (s == 2)
    hipLaunchKernelGGL((bulk_curv2d_multi<R, R, 2, TW, PL, WIDE>), grid, dim3(TW), 0, st, d, bds, wk, nw, mt, jac,
                       par, dfm);
  else if (s == 3)
    hipLaunchKernelGGL((bulk_curv2d_multi<R, R, 3, TW, PL, WIDE>), grid, dim3(TW), 0, st, d, bds, wk, nw, mt, jac,
                       par, dfm);
  else
    hipLaunchKernelGGL((bulk_curv2d_multi<R, R, 4, TW, PL, WIDE>), grid, dim3(TW), 0, st, d, bds, wk, nw, mt, jac,
                       par, dfm);
}

template <typename R, int TW, bool WIDE = false>
static void launch_mcurv_tw(const Tiled2D* d, const BlockDev* bds, const int2* wk, int nw, const R* mt, const R* jac,
                            int s, bool plastic, int par, int dfm, dim3 grid, hipStream_t st) {
  if (plastic) launch_mcurv_pl<R, TW, true, WIDE>(d, bds, wk, nw, mt, jac, s, par, dfm, grid, st);
  else launch_mcurv_pl<R, TW, false, WIDE>(d, bds, wk, nw, mt, jac, s, par, dfm, grid, st);
}

// stage descriptor of a 2-D rectilinear block: RK coefficients and arrays
template <typename R>
Tiled2D Solver<R>::stage_tile2(int b, const StageDev& sd) const {
  Tiled2D t2 = tiles2[b];
  t2.A = sd.A;
  t2.B = sd.B;
  t2.fA = (float)sd.A;
  t2.fB = (float)sd.B;
  t2.roundtrip = sd.roundtrip;
  t2.dfmode = sd.dfmode;
  t2.f[0] = f[0].p;
  t2.f[1] = f[1].p;
  t2.df = df.p;
  t2.hs = hs.p;
  return t2;
}

template <typename R>
void Solver<R>::launch_tiled2(Tiled2D t2, int tw, int par, hipStream_t st) {
  t2.nyt = (t2.hi[1] - t2.lo[1] + tw - 1) / tw;
  t2.ntiles = t2.nyt * t2.nxc;
  if (t2.wide) {
    if (tw != 64) throw Error{"2-D tiles with one descriptor per field: 64-wide tiles only"};
    launch_tiled2_tw<R, 64, true>(t2, S.sbporder, par, st);
  } else if (tw == 64) launch_tiled2_tw<R, 64>(t2, S.sbporder, par, st);
  else if (tw == 128) launch_tiled2_tw<R, 128>(t2, S.sbporder, par, st);
  else launch_tiled2_tw<R, 256>(t2, S.sbporder, par, st);
}

template <typename R>
void Solver<R>::launch_tiled(Tiled3D ta, int ty) {
  if (ta.rows_nht > 0) {
    launch_rows<R>(ta, S.sbporder, f[cur].p, f[cur ^ 1].p, df.p, hs.p, stream);
    return;
  }
  if (ta.wide) {  // one descriptor per field: the default tile height only
    constexpr int TYW = sizeof(R) == 8 ? 4 : 8;
    ta.nyt = (ta.hi[1] - ta.lo[1] + TYW - 1) / TYW;
    ta.ntiles = ta.nzt * ta.nyt * ta.nxc;
    launch_tiled_ty<R, TYW, true>(ta, S.sbporder, f[cur].p, f[cur ^ 1].p, df.p, hs.p, stream);
    return;
  }
  ta.nyt = (ta.hi[1] - ta.lo[1] + ty - 1) / ty;
  ta.ntiles = ta.nzt * ta.nyt * ta.nxc;
  if (ty == 2) launch_tiled_ty<R, 2>(ta, S.sbporder, f[cur].p, f[cur ^ 1].p, df.p, hs.p, stream);
  else if (ty == 4) launch_tiled_ty<R, 4>(ta, S.sbporder, f[cur].p, f[cur ^ 1].p, df.p, hs.p, stream);
  else launch_tiled_ty<R, 8>(ta, S.sbporder, f[cur].p, f[cur ^ 1].p, df.p, hs.p, stream);
}

template <typename R>
void Solver<R>::launch_bulk(int b, const StageDev& sd) {
  const BlockDev& D = bdev[b];
  const bool diss = S.cdiss > 0., het = S.hetmat;
  const bool tiled = tiled_ok[b];
  if (merged) {
    const Tiled2D* dsc = mdesc[(size_t)cur_stage * 2 + (sd.roundtrip ? 1 : 0)].p;
    const dim3 grid((unsigned)(((mnwork + 7) / 8) * 8));
    const int tw = tiles2[0].tw, s = S.sbporder;
    const int dm = sd.dfmode;
    if (tiles2[0].wide) {  // one descriptor per field (64-wide tiles)
      if (tw != 64) throw Error{"2-D tiles with one descriptor per field: 64-wide tiles only"};
      if (merged == 1) launch_merged_tw<R, 64, true>(dsc, mwork.p, mnwork, s, cur, dm, grid, stream);
      else launch_mcurv_tw<R, 64, true>(dsc, mbd.p, mwork.p, mnwork, metric.p, jac.p, s, S.plastic, cur, dm, grid,
                                        stream);
    } else if (merged == 1) {
      if (tw == 64) launch_merged_tw<R, 64>(dsc, mwork.p, mnwork, s, cur, dm, grid, stream);
      else if (tw == 128) launch_merged_tw<R, 128>(dsc, mwork.p, mnwork, s, cur, dm, grid, stream);
      else launch_merged_tw<R, 256>(dsc, mwork.p, mnwork, s, cur, dm, grid, stream);
    } else {
      if (tw == 64)
        launch_mcurv_tw<R, 64>(dsc, mbd.p, mwork.p, mnwork, metric.p, jac.p, s, S.plastic, cur, dm, grid, stream);
      else if (tw == 128)
        launch_mcurv_tw<R, 128>(dsc, mbd.p, mwork.p, mnwork, metric.p, jac.p, s, S.plastic, cur, dm, grid, stream);
      else launch_mcurv_tw<R, 256>(dsc, mbd.p, mwork.p, mnwork, metric.p, jac.p, s, S.plastic, cur, dm, grid, stream);
    }
    HIPCHK(hipGetLastError());
    return;
  }
  if (tiled && S.ndim == 2) {
    Tiled2D t2 = stage_tile2(b, sd);
    launch_tiled2(t2, t2.tw, cur, stream);
    HIPCHK(hipGetLastError());
    return;
  }
  if (curv_ok[b]) {
    Tiled2D t2 = stage_tile2(b, sd);
    t2.nyt = (t2.hi[1] - t2.lo[1] + t2.tw - 1) / t2.tw;
    t2.ntiles = t2.nyt * t2.nxc;
    if (t2.wide) {
      if (t2.tw != 64) throw Error{"2-D tiles with one descriptor per field: 64-wide tiles only"};
      launch_curv2_tw<R, 64, true>(t2, D, metric.p, jac.p, S.sbporder, S.plastic, cur, stream);
    } else if (t2.tw == 64) launch_curv2_tw<R, 64>(t2, D, metric.p, jac.p, S.sbporder, S.plastic, cur, stream);
    else if (t2.tw == 128) launch_curv2_tw<R, 128>(t2, D, metric.p, jac.p, S.sbporder, S.plastic, cur, stream);
    else launch_curv2_tw<R, 256>(t2, D, metric.p, jac.p, S.sbporder, S.plastic, cur, stream);
    HIPCHK(hipGetLastError());
    return;
  }
  if (tiled) {
    Tiled3D ta = tiles[b];
    ta.A = sd.A;
    ta.B = sd.B;
    ta.fA = (float)sd.A;
    ta.fB = (float)sd.B;
    ta.roundtrip = sd.roundtrip;
    ta.dfmode = sd.dfmode;
    launch_tiled(ta, tile_ty);
    HIPCHK(hipGetLastError());
    return;
  }
  const Boxes bx = make_boxes(D, false);
  switch (mode_id()) {
    case 0: bulk_dispatch<0, R>(diss, het, D, bx, G, sd, f[cur].p, f[cur ^ 1].p, df.p, metric.p, jac.p, mat.p, hs.p, stream); break;
    case 1: bulk_dispatch<1, R>(diss, het, D, bx, G, sd, f[cur].p, f[cur ^ 1].p, df.p, metric.p, jac.p, mat.p, hs.p, stream); break;
    default: bulk_dispatch<2, R>(diss, het, D, bx, G, sd, f[cur].p, f[cur ^ 1].p, df.p, metric.p, jac.p, mat.p, hs.p, stream); break;
  }
  HIPCHK(hipGetLastError());
}

// the merged 2-D stage launch with the interface solve in it
template <typename R>
void Solver<R>::launch_fused(const StageDev& sd, const FaultStageDev& fs) {
  const Tiled2D* dsc = mdesc[(size_t)cur_stage * 2 + (sd.roundtrip ? 1 : 0)].p;
  const dim3 grid((unsigned)(((mnfwork + 7) / 8) * 8));
  fdf::launch_fused<R>(tiles2[0].tw, dsc, mfwork.p, mnfwork, S.sbporder, cur, sd.dfmode, grid, mfdev.p, mpairs.p, fs,
                       err.p, stream);
  HIPCHK(hipGetLastError());
}

template <typename R>
void Solver<R>::launch_plastic(int b) {
  const BlockDev& D = bdev[b];
  const long long np = (long long)D.n[0] * D.n[1] * D.n[2];
  const dim3 grid((unsigned)((np + 255) / 256)), blk(256);
  switch (mode_id()) {
    case 0: hipLaunchKernelGGL((plastic_kernel<0, R>), grid, blk, 0, stream, D, G, S.dt, f[cur].p, fp.p, mat.p, hs.p); break;
    case 1: hipLaunchKernelGGL((plastic_kernel<1, R>), grid, blk, 0, stream, D, G, S.dt, f[cur].p, fp.p, mat.p, hs.p); break;
    default: hipLaunchKernelGGL((plastic_kernel<2, R>), grid, blk, 0, stream, D, G, S.dt, f[cur].p, fp.p, mat.p, hs.p); break;
  }
  HIPCHK(hipGetLastError());
}

template <typename R>
int Solver<R>::stage_dfmode(int st, bool full_step) const {
  if (opts.exact_fp64) return DF_ALL;  // the reference's scale_df and update, unchanged
  int m = F_STORE;
  if (S.rk.A[st] != 0.) m |= DF_LOAD;
  if (!(full_step && st + 1 == S.rk.nstages && S.rk.A[0] == 0.)) m |= DF_STORE;
  return m;
}

template <typename R>
void Solver<R>::materialize_df() {
  if (!df_stale) return;
  const int st = S.rk.nstages - 1;
  StageDev sd{S.rk.A[st], S.rk.B[st], S.dt, 0};
  sd.dfmode = (S.rk.A[st] != 0. ? DF_LOAD : 0) | DF_STORE;
  const int save_cur = cur, save_stage = cur_stage;
  cur ^= 1;  // f[cur] is again the last stage's input
  cur_stage = st;
  for (size_t b = 0; b < S.blocks.size(); b++) {
    if (merged && b > 0) break;
    launch_bulk((int)b, sd);
  }
  cur = save_cur;
  cur_stage = save_stage;
  HIPCHK(hipStreamSynchronize(stream));
  df_stale = false;
}

// domain::do_rk_stage, domain.cpp:228-305
template <typename R>
void Solver<R>::do_stage(int istep, int st, bool full_step) {
  const bool last = (st + 1 == S.rk.nstages);
  const int dfm = stage_dfmode(st, full_step);
  if (!capturing && df_stale && (dfm & DF_LOAD)) materialize_df();
  const double t = (double)istep * S.dt;
  const double tst = t + S.rk.C[st] * S.dt;
  const double tfront = (double)(istep + 1) * S.dt;
  FaultStageDev fs{S.rk.A[st], S.rk.B[st], S.dt, tst, 0, (full_step && last && !S.plastic) ? tfront : -1.,
                   capturing ? graph_step_ptr : nullptr, S.rk.C[st], (full_step && last && !S.plastic) ? 1 : 0};
  if (!capturing && !S.ifaces.empty() && !fused2d) {
    if (nev == NEV) flush_events();
    HIPCHK(hipEventRecord(evs[2 * nev], stream));
    evkind[nev] = 1;
  }
  if (!fused2d)
    for (size_t i = 0; i < S.ifaces.size(); i++) launch_fault((int)i, fs);
  if (!capturing && !S.ifaces.empty() && !fused2d) {
    HIPCHK(hipEventRecord(evs[2 * nev + 1], stream));
    nev++;
  }
  StageDev sd{S.rk.A[st], S.rk.B[st], S.dt, (full_step && last && !S.plastic && sizeof(R) == 8) ? 1 : 0, dfm};
  cur_stage = st;
  if (pair3) {  // the second block on the side stream, after the fault solve
    HIPCHK(hipEventRecord(ev_fork, stream));
    HIPCHK(hipStreamWaitEvent(side, ev_fork, 0));
  }
  for (size_t b = 0; b < S.blocks.size(); b++) {
    if (merged && b > 0) break;  // block 0's launch covers every block
    const hipStream_t main = stream;
    if (pair3 && b == 1) stream = side;  // launch_bulk issues on `stream`
    if (capturing) {
      if (fused2d) launch_fused(sd, fs);
      else launch_bulk((int)b, sd);
    } else {
      if (nev == NEV) flush_events();
      HIPCHK(hipEventRecord(evs[2 * nev], stream));
      evkind[nev] = 0;
      if (fused2d) launch_fused(sd, fs);
      else launch_bulk((int)b, sd);
      HIPCHK(hipEventRecord(evs[2 * nev + 1], stream));
      nev++;
    }
    stream = main;
  }
  if (pair3) {
    HIPCHK(hipEventRecord(ev_join, side));
    HIPCHK(hipStreamWaitEvent(stream, ev_join, 0));
  }
  cur ^= 1;
  if (!capturing) df_stale = !(dfm & DF_STORE);
  if (last && S.plastic) {
    for (size_t b = 0; b < S.blocks.size(); b++) launch_plastic((int)b);
    FaultStageDev fs2{0., 0., S.dt, tst, 1, full_step ? tfront : -1., capturing ? graph_step_ptr : nullptr, S.rk.C[st],
                      full_step ? 1 : 0};
    for (size_t i = 0; i < S.ifaces.size(); i++)
      if (S.ifaces[i].law != LAW_LOCKED) launch_fault((int)i, fs2);
  }
}

template <typename R>
void Solver<R>::flush_events() {
  for (int e = 0; e < nev; e++) {
    HIPCHK(hipEventSynchronize(evs[2 * e + 1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, evs[2 * e], evs[2 * e + 1]));
    if (evkind[e]) {
      fault_time_ms += ms;
      fault_count += 1;
    } else {
      bulk_time_ms += ms;
      bulk_count += 1;
    }
  }
  nev = 0;
}

// one time step captured as a hipGraph (one graph per f-buffer parity);
// times come from the device step counter so the graph replays unchanged
template <typename R>
void Solver<R>::graph_step() {
  const int par = cur;
  if (!graphs[par]) {
    HIPCHK(hipStreamSynchronize(stream));
    graph_step_ptr = dstep.p;
    capturing = true;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int st = 0; st < S.rk.nstages; st++) do_stage(cur_step, st, true);
    hipLaunchKernelGGL(advance_step, dim3(1), dim3(1), 0, stream, dstep.p);
    hipGraph_t g;
    HIPCHK(hipStreamEndCapture(stream, &g));
    capturing = false;
    HIPCHK(hipGraphInstantiate(&graphs[par], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    cur = par;  // capture did not execute; replay below flips the buffers
    graph_flip = (S.rk.nstages & 1) ? 1 : 0;
  }
  if (!dstep_valid) {
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(dstep.p, &cur_step, sizeof(int), hipMemcpyHostToDevice));
    dstep_valid = true;
  }
  HIPCHK(hipGraphLaunch(graphs[par], stream));
  cur ^= graph_flip;
  df_stale = !(stage_dfmode(S.rk.nstages - 1, true) & DF_STORE);
}

template <typename R>
void Solver<R>::step(int n) {
  for (int k = 0; k < n; k++) {
    if (opts.use_graph) {
      graph_step();
    } else {
      dstep_valid = false;
      for (int st = 0; st < S.rk.nstages; st++) do_stage(cur_step, st, true);
    }
    cur_step++;
    if (opts.write_outputs) write_outputs(cur_step);
    if (!opts.quiet && S.ninfo > 0 && cur_step % S.ninfo == 0) {
      time_t raw;
      time(&raw);
      printf("Timestep %d of %d %s", cur_step, S.nt, asctime(localtime(&raw)));
      fflush(stdout);
    }
  }
  sync_check();
}

// staging rings of the output units: as many slots of a unit as fit 64 MB,
// 2 to 16 (FDF_OUTPUT_SLOTS=k caps them at k)
template <typename R>
void Solver<R>::setup_output_staging() {
  HIPCHK(hipStreamCreateWithFlags(&ostream, hipStreamNonBlocking));
  ostage.resize(S.outputs.size());
  std::vector<OutGatherDev> gd(S.outputs.size());
  for (size_t u = 0; u < S.outputs.size(); u++) {
    const OutputH& U = S.outputs[u];
    OutStage& O = ostage[u];
    O.cnt = (size_t)U.nx[0] * U.nx[1] * U.nx[2];
    if (!O.cnt) continue;
    O.nslot = (int)std::max<size_t>(2, std::min<size_t>(16, ((size_t)64 << 20) / (O.cnt * sizeof(double))));
    if (const char* e = getenv("FDF_OUTPUT_SLOTS"))  // fewer slots (tests of a full ring)
      if (atoi(e) >= 1) O.nslot = std::min(O.nslot, atoi(e));
    O.dev.alloc(O.cnt * O.nslot);
    HIPCHK(hipHostMalloc((void**)&O.host, O.cnt * O.nslot * sizeof(double), hipHostMallocDefault));
    O.gathered.resize(O.nslot);
    O.copied.resize(O.nslot);
    O.busy.assign(O.nslot, 0);
    for (int k = 0; k < O.nslot; k++) {
      HIPCHK(hipEventCreateWithFlags(&O.gathered[k], hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&O.copied[k], hipEventDisableTiming));
    }
    OutGatherDev& D = gd[u];
    for (int d = 0; d < 3; d++) {
      D.lo[d] = U.xm[d];
      D.st[d] = U.xs[d];
      D.n[d] = U.nx[d];
    }
    D.dst = O.dev.p;
    D.cnt = (long long)O.cnt;
    const int hi[3] = {U.xm[0] + (U.nx[0] - 1) * U.xs[0], U.xm[1] + (U.nx[1] - 1) * U.xs[1],
                       U.xm[2] + (U.nx[2] - 1) * U.xs[2]};
    if (U.iface >= 0) {  // every index of the box inside the interface arrays
      const IfaceH& I = S.ifaces[U.iface];
      if (I.law == LAW_LOCKED) throw Error{"read_iface: interface is not frictional"};
      if ((U.fidx == 7 || U.fidx == 10) && I.law != LAW_STZ) throw Error{"read_iface: interface has no state"};
      const int a0 = (I.dir == 0) ? 1 : 0, a1 = (I.dir == 2) ? 1 : 2;
      const long n = (long)I.n0 * I.n1;
      const long ncomp = (long)iface_buf(U.iface, U.fidx)->n / std::max(n, 1L);
      if (U.xm[a0] < I.xm[a0] || hi[a0] - I.xm[a0] >= I.n0 || U.xm[a1] < I.xm[a1] || hi[a1] - I.xm[a1] >= I.n1 ||
          U.comp < 0 || U.comp >= ncomp)
        throw Error{"output unit " + U.name + ": index box outside the interface"};
      D.iface = 1;
      D.arr = iface_buf(U.iface, U.fidx)->p;
      D.n_if = n;
      D.nn1 = I.n1;
      D.comp = U.comp;
      D.a0 = a0;
      D.a1 = a1;
      D.xa0 = I.xm[a0];
      D.xa1 = I.xm[a1];
    } else {
      for (int d = 0; d < 3; d++)
        if (U.xm[d] < 0 || hi[d] >= S.nx[d]) throw Error{"output unit " + U.name + ": index box outside the grid"};
      if (U.fidx < 0 || U.fidx >= S.nf + S.nfp) throw Error{"output unit " + U.name + ": bad field"};
      D.field = U.fidx;
      D.hsf = -1;
      if (U.fidx >= S.nv && U.fidx < S.nv + S.ns) {
        const int i = U.fidx - S.nv;
        D.add = 1;
        D.s0add = S.s0[S.sidx[i]];
        if (S.hetstress) D.hsf = i;
      }
    }
  }
  ogather.upload(gd);
}

// write the oldest pending output slot once its copy has landed
template <typename R>
bool Solver<R>::drain_output(bool wait) {
  if (opend.empty()) return false;
  const int u = opend.front().first, k = opend.front().second;
  OutStage& O = ostage[u];
  if (!wait) {
    const hipError_t q = hipEventQuery(O.copied[k]);
    if (q == hipErrorNotReady) {
      (void)hipGetLastError();  // "not ready" is no error: keep it out of the next launch check
      return false;
    }
    HIPCHK(q);
  }
  HIPCHK(hipEventSynchronize(O.copied[k]));
  fwrite(O.host + (size_t)k * O.cnt, sizeof(double), O.cnt, outs[u].data);
  O.busy[k] = 0;
  opend.pop_front();
  return true;
}

template <typename R>
void Solver<R>::drain_outputs() {
  while (drain_output(true)) {
  }
}

template <typename R>
DBuf<double>* Solver<R>::iface_buf(int iface, int which) {
  if (iface < 0 || iface >= (int)S.ifaces.size()) throw Error{"read_iface: bad interface"};
  IfBufs& ib = ibufs[iface];
  switch (which) {
    case 0: return &ib.u;
    case 1: return &ib.v;
    case 2: return &ib.s;
    case 3: return &ib.sn;
    case 4: return &ib.ux;
    case 5: return &ib.vx;
    case 6: return &ib.sx;
    case 7: return &ib.state;
    case 8: return &ib.du;
    case 9: return &ib.dux;
    case 10: return &ib.dstate;
    default: throw Error{"read_iface: bad array id"};
  }
}

// outputunit::write_unit, outputunit.cpp:771-889: stress fields absolute
// (f + s0 + s, the order of fields::set_stress), interface arrays by
// component; gathered on stream st (default: the solver's own), which must
// hold the step's last launch
template <typename R>
void Solver<R>::write_outputs(int tstep, hipStream_t st) {
  if (!st) st = stream;
  OutGatherList L{};
  long long most = 0;
  auto launch = [&]() {
    if (!L.n) return;
    const dim3 grid((unsigned)((most + 255) / 256), (unsigned)L.n);
    hipLaunchKernelGGL((gather_units<R>), grid, dim3(256), 0, st, ogather.p, L, f[cur].p, fp.p, S.nf, hs.p, G.N,
                       G.Ny, G.Nz);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < L.n; q++) {  // copy each unit's slot once the gather is done
      OutStage& O = ostage[L.unit[q]];
      const int k = L.slot[q];
      if (q == 0) {
        HIPCHK(hipEventRecord(O.gathered[k], st));
        HIPCHK(hipStreamWaitEvent(ostream, O.gathered[k], 0));
      }
      HIPCHK(hipMemcpyAsync(O.host + (size_t)k * O.cnt, O.dev.p + (size_t)k * O.cnt, O.cnt * sizeof(double),
                            hipMemcpyDeviceToHost, ostream));
      HIPCHK(hipEventRecord(O.copied[k], ostream));
      opend.emplace_back(L.unit[q], k);
    }
    L.n = 0;
    most = 0;
  };
  for (size_t u = 0; u < S.outputs.size(); u++) {
    const OutputH& U = S.outputs[u];
    if (tstep < U.tm || tstep > U.tp || (tstep - U.tm) % U.ts != 0) continue;
    if (U.tp > U.tm) {
      const double t = (double)tstep * S.dt;
      fwrite(&t, sizeof(double), 1, outs[u].time);
    }
    OutStage& O = ostage[u];
    if (!O.cnt) continue;
    const int k = O.next;
    O.next = (O.next + 1) % O.nslot;
    while (O.busy[k]) {  // the ring is full: write the oldest pending slots
      if (opend.empty()) launch();  // (this step's own earlier units hold it)
      drain_output(true);
    }
    O.busy[k] = 1;
    L.unit[L.n] = (int)u;
    L.slot[L.n] = k;
    L.n++;
    most = std::max(most, (long long)O.cnt);
    if (L.n == OutGatherList::MAXU) launch();
  }
  launch();
  while (drain_output(false)) {
  }
}

template <typename R>
void Solver<R>::coords_box(int c, const int lo[3], const int hi[3], const int st[3], std::vector<double>& out) {
  int n[3];
  for (int d = 0; d < 3; d++) n[d] = hi[d] < lo[d] ? 0 : (hi[d] - lo[d]) / st[d] + 1;
  const size_t cnt = (size_t)n[0] * n[1] * n[2];
  out.clear();
  if (!cnt) return;
  if (!dgrid) {
    out.reserve(cnt);
    for (int a = lo[0]; a <= hi[0]; a += st[0])
      for (int b = lo[1]; b <= hi[1]; b += st[1])
        for (int cc = lo[2]; cc <= hi[2]; cc += st[2]) out.push_back(S.x[(size_t)c * S.N + S.idx(a, b, cc)]);
    return;
  }
  out.resize(cnt);
  DBuf<double> tmp;
  tmp.alloc(cnt);
  hipLaunchKernelGGL((gather_box<double>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream,
                     x.p + (size_t)c * S.N, hs.p, 0., -1, G.N, G.Ny, G.Nz, lo[0], lo[1], lo[2], st[0], st[1], st[2],
                     n[0], n[1], n[2], tmp.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(out.data(), tmp.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
}

template <typename R>
void Solver<R>::read_field(int field, const int lo[3], const int hi[3], const int st[3], double* h) {
  int n[3];
  for (int d = 0; d < 3; d++) {
    if (lo[d] < 0 || hi[d] >= S.nx[d] || hi[d] < lo[d] || st[d] <= 0) throw Error{"read_field: bad index box"};
    n[d] = (hi[d] - lo[d]) / st[d] + 1;
  }
  const size_t cnt = (size_t)n[0] * n[1] * n[2];
  const R* src;
  int q = field;
  if (field < S.nf) src = f[cur].p + (size_t)field * S.N;
  else if (field < S.nf + S.nfp) { src = fp.p + (size_t)(field - S.nf) * S.N; q = -1; }
  else throw Error{"read_field: bad field"};
  (void)q;
  DBuf<double> tmp;
  tmp.alloc(cnt);
  const unsigned nb = (unsigned)((cnt + 255) / 256);
  hipLaunchKernelGGL((gather_box<R>), dim3(nb), dim3(256), 0, stream, src, hs.p, 0., -1, G.N, G.Ny, G.Nz, lo[0], lo[1],
                     lo[2], st[0], st[1], st[2], n[0], n[1], n[2], tmp.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(h, tmp.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
}

template <typename R>
void Solver<R>::read_array(int which, double* h, long n) {
  HIPCHK(hipStreamSynchronize(stream));
  const size_t N = S.N;
  auto fetch = [&](const R* src, size_t cnt, double* dst) {
    if (std::is_same<R, double>::value) {
      HIPCHK(hipMemcpy(dst, src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<R> tmp(cnt);
      HIPCHK(hipMemcpy(tmp.data(), src, cnt * sizeof(R), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < cnt; i++) dst[i] = (double)tmp[i];
    }
  };
  switch (which) {
    case 0: {
      const size_t need = (size_t)(S.nf + S.nfp) * N;
      if ((size_t)n != need) throw Error{"read_array: size mismatch"};
      fetch(f[cur].p, (size_t)S.nf * N, h);
      if (S.nfp) fetch(fp.p, (size_t)S.nfp * N, h + (size_t)S.nf * N);
      break;
    }
    case 1:
      if ((size_t)n != (size_t)S.nf * N) throw Error{"read_array: size mismatch"};
      materialize_df();
      fetch(df.p, (size_t)n, h);
      break;
    case 2:
    case 3:
    case 4: {
      const size_t want = which == 2 ? (size_t)S.ndim * N : (which == 3 ? (size_t)S.ndim * S.ndim * N : N);
      if ((size_t)n != want) throw Error{"read_array: size mismatch"};
      if (dgrid) {
        const double* src = which == 2 ? x.p : (which == 3 ? metric64 : jac64);
        HIPCHK(hipMemcpy(h, src, want * sizeof(double), hipMemcpyDeviceToHost));
      } else {
        const double* src = which == 2 ? S.x.data() : (which == 3 ? S.metric.data() : S.jac.data());
        memcpy(h, src, want * sizeof(double));
      }
      break;
    }
    default:
      throw Error{"read_array: bad array id"};
  }
}

template <typename R>
void Solver<R>::write_array(int which, const double* h, long n) {
  HIPCHK(hipStreamSynchronize(stream));
  const size_t N = S.N;
  auto put = [&](R* dst, const double* src, size_t cnt) {
    std::vector<R> tmp(cnt);
    for (size_t i = 0; i < cnt; i++) tmp[i] = (R)src[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), cnt * sizeof(R), hipMemcpyHostToDevice));
  };
  if (which == 0) {
    if ((size_t)n != (size_t)(S.nf + S.nfp) * N) throw Error{"write_array: size mismatch"};
    put(f[cur].p, h, (size_t)S.nf * N);
    if (S.nfp) put(fp.p, h + (size_t)S.nf * N, (size_t)S.nfp * N);
  } else if (which == 1) {
    if ((size_t)n != (size_t)S.nf * N) throw Error{"write_array: size mismatch"};
    put(df.p, h, (size_t)n);
    df_stale = false;
  } else {
    throw Error{"write_array: only f and df are writable"};
  }
}

template <typename R>
long Solver<R>::read_iface(int iface, int which, double* h, long n) {
  DBuf<double>* b = iface_buf(iface, which);
  if (S.ifaces[iface].law == LAW_LOCKED) throw Error{"read_iface: interface is not frictional"};
  if ((which == 7 || which == 10) && S.ifaces[iface].law != LAW_STZ) throw Error{"read_iface: interface has no state"};
  if (!h) return (long)b->n;
  if (n < (long)b->n) throw Error{"read_iface: buffer too small"};
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(h, b->p, b->n * sizeof(double), hipMemcpyDeviceToHost));
  return (long)b->n;
}

template <typename R>
long Solver<R>::read_front(int iface, double* h, long n) {
  if (iface < 0 || iface >= (int)S.ifaces.size() || !ibufs[iface].front.p) return -1;
  DBuf<double>& b = ibufs[iface].front;
  if (!h) return (long)b.n;
  if (n < (long)b.n) throw Error{"read_front: buffer too small"};
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(h, b.p, b.n * sizeof(double), hipMemcpyDeviceToHost));
  return (long)b.n;
}

// State snapshot for a later resume (the reference has no restart): a
// header (tag, version, bulk word size, N, nf, nfp, interface count, step),
// then f, df and the plastic fields in the bulk precision, then per
// interface its slip/state arrays and front in FP64.  The interface SAT
// terms are not state: every stage's fault solve writes them before the bulk
// update reads them.
static const char kStateTag[8] = {'F', 'D', 'F', 'S', 'T', 'A', 'T', 'E'};
static const int kStateVersion = 1;
static const int kStateIface[] = {0, 8, 1, 2, 3, 4, 9, 5, 6, 7, 10};  // u du v s sn ux dux vx sx state dstate

template <typename R>
void Solver<R>::save_state(const char* path) {
  materialize_df();
  HIPCHK(hipStreamSynchronize(stream));
  FILE* fh = fopen(path, "wb");
  if (!fh) throw Error{std::string("save_state: cannot open ") + path};
  std::unique_ptr<FILE, int (*)(FILE*)> guard(fh, fclose);
  auto put = [&](const void* p, size_t bytes) {
    if (fwrite(p, 1, bytes, fh) != bytes) throw Error{std::string("save_state: cannot write ") + path};
  };
  auto put_dev = [&](const void* d, size_t bytes) {
    std::vector<char> h(bytes);
    if (bytes) HIPCHK(hipMemcpy(h.data(), d, bytes, hipMemcpyDeviceToHost));
    put(h.data(), bytes);
  };
  const long long hdr[7] = {kStateVersion, (long long)sizeof(R), (long long)S.N, S.nf, S.nfp,
                            (long long)S.ifaces.size(), cur_step};
  put(kStateTag, sizeof(kStateTag));
  put(hdr, sizeof(hdr));
  put_dev(f[cur].p, (size_t)S.nf * S.N * sizeof(R));
  put_dev(df.p, (size_t)S.nf * S.N * sizeof(R));
  if (S.nfp) put_dev(fp.p, (size_t)S.nfp * S.N * sizeof(R));
  for (size_t i = 0; i < ibufs.size(); i++) {
    IfBufs& ib = ibufs[i];
    for (int w : kStateIface) {
      DBuf<double>* b = iface_buf((int)i, w);
      const long long n = (long long)b->n;
      put(&n, sizeof(n));
      put_dev(b->p, b->n * sizeof(double));
    }
    const long long nf_ = (long long)ib.front.n;
    put(&nf_, sizeof(nf_));
    put_dev(ib.front.p, ib.front.n * sizeof(double));
  }
  // a write that fails only when the buffered data is flushed shows at fclose
  if (fclose(guard.release()) != 0) throw Error{std::string("save_state: cannot write ") + path};
}

template <typename R>
void Solver<R>::load_state(const char* path) {
  FILE* fh = fopen(path, "rb");
  if (!fh) throw Error{std::string("load_state: cannot open ") + path};
  std::unique_ptr<FILE, int (*)(FILE*)> guard(fh, fclose);
  auto get = [&](void* p, size_t bytes) {
    if (fread(p, 1, bytes, fh) != bytes) throw Error{std::string("load_state: truncated snapshot ") + path};
  };
  char tag[8];
  long long hdr[7];
  get(tag, sizeof(tag));
  get(hdr, sizeof(hdr));
  if (memcmp(tag, kStateTag, sizeof(tag)) != 0 || hdr[0] != kStateVersion)
    throw Error{std::string("load_state: not a state snapshot: ") + path};
  if (hdr[1] != (long long)sizeof(R) || hdr[2] != (long long)S.N || hdr[3] != S.nf || hdr[4] != S.nfp ||
      hdr[5] != (long long)S.ifaces.size())
    throw Error{"load_state: the snapshot is of another problem or precision"};
  if (hdr[6] < 0 || hdr[6] > INT_MAX) throw Error{std::string("load_state: not a state snapshot: ") + path};
  // read everything into host memory first, so a bad file leaves the state as it was
  std::vector<char> hf((size_t)S.nf * S.N * sizeof(R)), hdf(hf.size()), hfp((size_t)S.nfp * S.N * sizeof(R));
  get(hf.data(), hf.size());
  get(hdf.data(), hdf.size());
  if (S.nfp) get(hfp.data(), hfp.size());
  std::vector<std::vector<double>> hif;
  for (size_t i = 0; i < ibufs.size(); i++) {
    for (int w : kStateIface) {
      long long n = 0;
      get(&n, sizeof(n));
      if (n != (long long)iface_buf((int)i, w)->n) throw Error{"load_state: the snapshot is of another problem"};
      hif.emplace_back((size_t)n);
      get(hif.back().data(), (size_t)n * sizeof(double));
    }
    long long n = 0;
    get(&n, sizeof(n));
    if (n != (long long)ibufs[i].front.n) throw Error{"load_state: the snapshot is of another problem"};
    hif.emplace_back((size_t)n);
    get(hif.back().data(), (size_t)n * sizeof(double));
  }
  char extra;
  if (fread(&extra, 1, 1, fh) != 0) throw Error{std::string("load_state: trailing data in ") + path};
  drain_outputs();
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(f[cur].p, hf.data(), hf.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(df.p, hdf.data(), hdf.size(), hipMemcpyHostToDevice));
  if (S.nfp) HIPCHK(hipMemcpy(fp.p, hfp.data(), hfp.size(), hipMemcpyHostToDevice));
  size_t k = 0;
  for (size_t i = 0; i < ibufs.size(); i++) {
    for (int w : kStateIface) {
      DBuf<double>* b = iface_buf((int)i, w);
      if (b->n) HIPCHK(hipMemcpy(b->p, hif[k].data(), b->n * sizeof(double), hipMemcpyHostToDevice));
      k++;
    }
    if (ibufs[i].front.n)
      HIPCHK(hipMemcpy(ibufs[i].front.p, hif[k].data(), ibufs[i].front.n * sizeof(double), hipMemcpyHostToDevice));
    k++;
  }
  cur_step = (int)hdr[6];
  dstep_valid = false;  // the device step counter follows at the next graph replay
  df_stale = false;
}

// outputlist::close_list + frontlist::write_list (front.cpp:127-359)
template <typename R>
void Solver<R>::finish() {
  drain_outputs();
  for (OutFile& o : outs) {
    if (o.data) fclose(o.data);
    if (o.time) fclose(o.time);
    o.data = o.time = nullptr;
  }
  if (!opts.write_outputs || !S.has_front) return;
  const char endian = [] {
    const uint32_t v = 0x01020304;
    unsigned char c[4];
    memcpy(c, &v, 4);
    return c[0] == 4 ? '<' : (c[0] == 1 ? '>' : '=');
  }();
  for (size_t i = 0; i < S.ifaces.size(); i++) {
    const IfaceH& I = S.ifaces[i];
    if (I.law == LAW_LOCKED) continue;
    const std::string pre = S.datadir + S.name + "_front_" + std::to_string(i);
    std::vector<double> tv((size_t)I.n0 * I.n1);
    read_front((int)i, tv.data(), (long)tv.size());
    FILE* fpp = fopen((pre + "_t.dat").c_str(), "wb");
    if (!fpp) throw Error{"Error opening file in front.cpp"};
    fwrite(tv.data(), sizeof(double), tv.size(), fpp);
    fclose(fpp);
    static const char* xyz[3] = {"x", "y", "z"};
    for (int c = 0; c < S.ndim; c++) {
      FILE* xf = fopen((pre + "_" + xyz[c] + ".dat").c_str(), "wb");
      if (!xf) throw Error{"Error opening file in outputunit.cpp"};
      std::vector<double> buf;
      const int one[3] = {1, 1, 1};
      coords_box(c, I.xm, I.xp, one, buf);
      fwrite(buf.data(), sizeof(double), buf.size(), xf);
      fclose(xf);
    }
    FILE* m = fopen((pre + ".m").c_str(), "w");
    if (!m) throw Error{"Error writing parameters to matlab file"};
    fprintf(m, "endian = '%c';\nnx = %d;\nny = %d;\n", endian == '<' ? 'l' : (endian == '>' ? 'b' : 'n'), I.n0, I.n1);
    fclose(m);
    FILE* py = fopen((pre + ".py").c_str(), "w");
    if (!py) throw Error{"Error writing parameters to python file"};
    fprintf(py, "endian = '%c'\nnx = %d\nny = %d\n", endian, I.n0, I.n1);
    fclose(py);
  }
}

template <typename R>
double Solver<R>::fault_ms(int* n) {
  flush_events();
  const double avg = fault_count ? fault_time_ms / fault_count : 0.;
  if (n) *n = fault_count;
  fault_time_ms = 0.;
  fault_count = 0;
  return avg;
}

template <typename R>
double Solver<R>::bulk_ms(int* n) {
  flush_events();
  const double avg = bulk_count ? bulk_time_ms / bulk_count : 0.;
  if (n) *n = bulk_count;
  bulk_time_ms = 0.;
  bulk_count = 0;
  return avg;
}

// ---------------------------------------------------------------------------
// Ensemble of independent 2-D mode II ruptures stepped together.  Each
// instance is set up exactly like a single run (parse, grid, t = 0 solve,
// outputs); the batch then replaces the per-instance stage launches by two
// launches per RK stage over every instance: fault_batch_kernel (all fault
// points) and bulk_tiled2d_batch (all tiles of all blocks), the latter either
// as a persistent grid walking the work list or with one workgroup per item.
// Stage descriptors are precomputed per RK stage, so one time step is a fixed
// launch sequence that replays from a hipGraph.  Tile width, x-chunk and
// launch mode come from a timing autotuner run at open (cached per device and
// ensemble).
class BatchBase {
 public:
  virtual ~BatchBase() {}
  virtual void step(int n) = 0;
  virtual SolverBase* instance(int i) = 0;
  virtual int size() const = 0;
  virtual void sync() = 0;
  virtual void finish() = 0;
  virtual int fused_instances() const { return 0; }  // instances whose fault solve is in the stage launch
  int tw = 0, cx = 0;
  int persist = 1;  // 1: persistent grid over the work list, 0: one workgroup per item
  double tune_ms = 0.;
  int tuned_from_cache = 0;
};

template <typename R>
class Batch : public BatchBase {
 public:
  Batch(const std::vector<std::string>& paths, const fdf_opts& o);
  ~Batch() override;
  void step(int n) override;
  SolverBase* instance(int i) override { return inst.at(i).get(); }
  int size() const override { return (int)inst.size(); }
  void sync() override;
  void finish() override;

 private:
  void build(int tw_, int cx_, int persist_);
  void launch_stage(int st);
  double time_config(int tw_, int cx_, int persist_);
  void autotune();
  template <int TW>
  void launch_bulk_tw(int st, int dfm);
  int stage_dfmode(int st) const;

  fdf_opts opts{};
  std::vector<std::unique_ptr<Solver<R>>> inst;
  hipStream_t stream = nullptr;
  int nst = 0, sbp = 0, rko = 0, cur = 0, cur_step = 0, nt = 0;
  LowStorageRK rk;
  std::vector<DBuf<Tiled2D>> descs;  // per RK stage, all blocks of all instances
  DBuf<int2> work;
  int nwork = 0, grid = 0;
  DBuf<FaultBatchDev> fdescs;
  DBuf<int2> fwork;
  int nfwork = 0;
  std::vector<int> inst_fd;  // fault descriptor of each instance's first interface
  int flaw = -1;             // the friction law every interface shares (-1: mixed)
  // every instance with one slip-weakening y-interface whose tiles line up:
  // one launch per stage (bulk_fault2d_batch) over the tiles and the fault
  // chunks of all instances (off by default; FDF_FUSE2D=1 forces it)
  bool fused = false;
  DBuf<int2> ffwork;
  DBuf<FaultPair2D> fpairs;
  int nffwork = 0, ffgrid = 0;

 public:
  int fused_instances() const override { return fused ? (int)inst.size() : 0; }

 private:
  DBuf<int> err, dstep;
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  bool dstep_valid = false;
};

template <typename R>
Batch<R>::Batch(const std::vector<std::string>& paths, const fdf_opts& o) : opts(o) {
  if (paths.empty()) throw Error{"fdf_open_batch: no instances"};
  for (size_t i = 0; i < paths.size(); i++) {
    fdf_opts oi = o;
    oi.use_graph = 0;
    inst.emplace_back(new Solver<R>(paths[i], oi));
    const Solver<R>& I = *inst.back();
    const SetupH& S = I.S;
    if (S.ndim != 2 || S.mode != 2 || S.plastic)
      throw Error{"fdf_open_batch: instance " + std::to_string(i) + " is not a 2-D mode II elastic problem"};
    for (size_t b = 0; b < S.blocks.size(); b++) {
      if (!I.tiled_ok[b])
        throw Error{"fdf_open_batch: instance " + std::to_string(i) +
                    " has a block that is not rectilinear and homogeneous (or uses dissipation)"};
      if (I.tiles2[b].wide)
        throw Error{"fdf_open_batch: instance " + std::to_string(i) +
                    " has field arrays above 4 GiB (open it on its own with fdf_open)"};
    }
    if (i == 0) {
      nst = S.rk.nstages;
      sbp = S.sbporder;
      rko = S.rkorder;
      rk = S.rk;
      nt = S.nt;
    } else if (S.rkorder != rko || S.sbporder != sbp) {
      throw Error{"fdf_open_batch: instances must share sbporder and rkorder"};
    }
  }
  HIPCHK(hipSetDevice(o.device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto& I : inst) HIPCHK(hipStreamSynchronize(I->stream));
  err.alloc(1);
  err.zero(stream);
  dstep.alloc(1);
  dstep.zero(stream);
  // fault work list: (instance interface, first point of a chunk of 256)
  std::vector<FaultBatchDev> fd;
  std::vector<int2> fw;
  for (auto& I : inst) {
    inst_fd.push_back(I->fdev.empty() ? -1 : (int)fd.size());
    for (size_t k = 0; k < I->fdev.size(); k++) {
      FaultBatchDev d{};
      d.F = I->fdev[k];
      d.g = I->G;
      d.f[0] = I->f[0].p;
      d.f[1] = I->f[1].p;
      d.mat = I->mat.p;
      d.hs = I->hs.p;
      d.dt = I->S.dt;
      for (int p0 = 0; p0 < d.F.n; p0 += 256) fw.push_back(make_int2((int)fd.size(), p0));
      fd.push_back(d);
    }
  }
  for (size_t k = 0; k < fd.size(); k++) flaw = (k == 0 || fd[k].F.law == flaw) ? fd[k].F.law : -2;
  if (flaw < 0) flaw = -1;
  fdescs.upload(fd);
  fwork.upload(fw);
  nfwork = (int)fw.size();
  descs.resize(nst);
  autotune();
  HIPCHK(hipStreamSynchronize(stream));
}

template <typename R>
Batch<R>::~Batch() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (int q = 0; q < 2; q++)
    if (graphs[q]) (void)hipGraphExecDestroy(graphs[q]);
  if (stream) (void)hipStreamDestroy(stream);
}

// per-stage descriptors and the tile work list for tile width tw_, chunk cx_
template <typename R>
void Batch<R>::build(int tw_, int cx_, int persist_) {
  tw = tw_;
  cx = cx_;
  persist = persist_;
  std::vector<int2> wl;
  std::vector<Tiled2D> geo;  // stage 0 descriptors: the tile geometry
  for (int st = 0; st < nst; st++) {
    std::vector<Tiled2D> d;
    for (auto& I : inst) {
      const StageDev sd{rk.A[st], rk.B[st], I->S.dt, (st + 1 == nst && sizeof(R) == 8) ? 1 : 0};
      for (size_t b = 0; b < I->S.blocks.size(); b++) {
        Tiled2D t2 = I->stage_tile2((int)b, sd);
        t2.cx = cx;
        t2.nyt = (t2.hi[1] - t2.lo[1] + tw - 1) / tw;
        t2.nxc = (t2.hi[0] - t2.lo[0] + cx - 1) / cx;
        t2.ntiles = t2.nyt * t2.nxc;
        if (st == 0)
          for (int tl = 0; tl < t2.ntiles; tl++) wl.push_back(make_int2((int)d.size(), tl));
        d.push_back(t2);
      }
    }
    if (st == 0) geo = d;
    descs[st].upload(d);
  }
  work.upload(wl);
  nwork = (int)wl.size();
  // the fused list: every instance's fault chunks, and its tiles apart from
  // the interface tiles that the chunks take
  {
    std::vector<FaultPair2D> pairs;
    std::vector<std::pair<int, int>> skip;  // (descriptor, y tile)
    // off by default: at 3 waves per SIMD the ensemble's large bulk stage
    // (the FP32 tiles run 6 on their own) loses more than the separate fault
    // launch costs (FDF_FUSE2D=1 forces it)
    const char* ef = getenv("FDF_FUSE2D");
    bool ok = ef && atoi(ef) != 0;
    int base = 0;
    for (size_t i = 0; i < inst.size() && ok; i++) {
      const Solver<R>& I = *inst[i];
      int yA = 0, yB = 0;
      ok = I.S.ifaces.size() == 1 && inst_fd[i] >= 0;
      if (ok) {
        const IfaceH& F = I.S.ifaces[0];
        ok = fault_chunks_2d(F, I.fdev[0], geo[base + F.b1], geo[base + F.b2], tw, inst_fd[i], base + F.b1,
                             base + F.b2, pairs, yA, yB);
        skip.emplace_back(base + F.b1, yA);
        skip.emplace_back(base + F.b2, yB);
      }
      base += (int)I.S.blocks.size();
    }
    fused = ok && !pairs.empty();
    if (fused) {
      std::vector<int2> tl;
      for (const int2& it : wl) {
        const int yt = it.y % geo[it.x].nyt;
        bool drop = false;
        for (const auto& sk : skip) drop = drop || (sk.first == it.x && sk.second == yt);
        if (!drop) tl.push_back(it);
      }
      const int nitems = (int)(tl.size() + pairs.size());
      fpairs.upload(pairs);
      int dev0 = 0, ncu0 = 0, occ0 = 0;
      HIPCHK(hipGetDevice(&dev0));
      HIPCHK(hipDeviceGetAttribute(&ncu0, hipDeviceAttributeMultiprocessorCount, dev0));
      if (tw == 64)
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ0, bulk_fault2d_batch<R, R, 3, 64, LAW_SLIPWEAK>, 64, 0));
      else if (tw == 128)
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ0, bulk_fault2d_batch<R, R, 3, 128, LAW_SLIPWEAK>, 128, 0));
      else
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ0, bulk_fault2d_batch<R, R, 3, 256, LAW_SLIPWEAK>, 256, 0));
      ffgrid = persist ? std::max(8, (std::min(nitems, ncu0 * std::max(occ0, 1)) / 8) * 8) : 0;
      const std::vector<int2> wlf = fused_work_list(tl, (int)pairs.size(), ffgrid);
      nffwork = (int)wlf.size();
      ffwork.upload(wlf);
    }
  }
  // persistent grid: as many workgroups as the device keeps resident (a
  // multiple of 8 for the XCD mapping); otherwise one per work item
  int dev = 0, ncu = 0, occ = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  if (tw == 64) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bulk_tiled2d_batch<R, R, 3, 64>, 64, 0));
  else if (tw == 128) HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bulk_tiled2d_batch<R, R, 3, 128>, 128, 0));
  else HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bulk_tiled2d_batch<R, R, 3, 256>, 256, 0));
  grid = persist ? std::max(8, (std::min(nwork, ncu * std::max(occ, 1)) / 8) * 8) : ((nwork + 7) / 8) * 8;
  for (int q = 0; q < 2; q++)
    if (graphs[q]) {
      HIPCHK(hipGraphExecDestroy(graphs[q]));
      graphs[q] = nullptr;
    }
}

template <typename R>
template <int TW>
void Batch<R>::launch_bulk_tw(int st, int dfm) {
  const Tiled2D* d = descs[st].p;
  const dim3 g(grid), b(TW);
  if (sbp == 2) hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 2, TW>), g, b, 0, stream, d, work.p, nwork, cur, dfm);
  else if (sbp == 3) hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 3, TW>), g, b, 0, stream, d, work.p, nwork, cur, dfm);
  else hipLaunchKernelGGL((bulk_tiled2d_batch<R, R, 4, TW>), g, b, 0, stream, d, work.p, nwork, cur, dfm);
}

// df traffic of stage st in a full step (as Solver::stage_dfmode; a batch
// always runs the tiled kernels, never the bit-faithful mode)
template <typename R>
int Batch<R>::stage_dfmode(int st) const {
  int m = F_STORE;
  if (rk.A[st] != 0.) m |= DF_LOAD;
  if (!(st + 1 == nst && rk.A[0] == 0.)) m |= DF_STORE;
  return m;
}

// one RK stage of every instance: the stage of domain::do_rk_stage
// (domain.cpp:228-305) with the fault update first, as in Solver::do_stage.
// Stage and front times come from the device step counter with each
// instance's own dt (t = step*dt + C*dt, front (step+1)*dt), the same
// expressions as a single run.
template <typename R>
void Batch<R>::launch_stage(int st) {
  const bool last = st + 1 == nst;
  const FaultStageDev fs{rk.A[st], rk.B[st], 0., 0., 0, -1., dstep.p, rk.C[st], last ? 1 : 0};
  if (fused) {
    fdf::launch_fused<R>(tw, descs[st].p, ffwork.p, nffwork, sbp, cur, stage_dfmode(st), dim3(ffgrid), fdescs.p,
                         fpairs.p, fs, err.p, stream);
    HIPCHK(hipGetLastError());
    cur ^= 1;
    return;
  }
  const int fgrid = std::max(1, std::min(nfwork, 4096));
  if (nfwork) {
    if (flaw == LAW_SLIPWEAK)
      hipLaunchKernelGGL((fault_batch_kernel<0, R, LAW_SLIPWEAK>), dim3(fgrid), dim3(256), 0, stream, fdescs.p, fwork.p,
                         nfwork, fs, cur, err.p);
    else
      hipLaunchKernelGGL((fault_batch_kernel<0, R>), dim3(fgrid), dim3(256), 0, stream, fdescs.p, fwork.p, nfwork, fs,
                         cur, err.p);
  }
  const int dfm = stage_dfmode(st);
  if (tw == 64) launch_bulk_tw<64>(st, dfm);
  else if (tw == 128) launch_bulk_tw<128>(st, dfm);
  else launch_bulk_tw<256>(st, dfm);
  HIPCHK(hipGetLastError());
  cur ^= 1;
}

// time of one stage's bulk launch for a configuration (inputs unchanged:
// df is saved and restored around the trial launches, f_new is overwritten
// by the next stage anyway)
template <typename R>
double Batch<R>::time_config(int tw_, int cx_, int persist_) {
  build(tw_, cx_, persist_);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int save = cur;
  // a middle stage's full traffic (df loaded and stored)
  if (tw == 64) launch_bulk_tw<64>(0, DF_ALL);
  else if (tw == 128) launch_bulk_tw<128>(0, DF_ALL);
  else launch_bulk_tw<256>(0, DF_ALL);
  const int reps = 3;
  HIPCHK(hipEventRecord(e0, stream));
  for (int k = 0; k < reps; k++) {
    if (tw == 64) launch_bulk_tw<64>(0, DF_ALL);
    else if (tw == 128) launch_bulk_tw<128>(0, DF_ALL);
    else launch_bulk_tw<256>(0, DF_ALL);
  }
  HIPCHK(hipEventRecord(e1, stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipGetLastError());
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  cur = save;
  return ms / reps;
}

template <typename R>
void Batch<R>::autotune() {
  // key: device, precision, order, and every instance's grid shape
  hipDeviceProp_t prop;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  unsigned long long h = 1469598103934665603ULL;
  auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ULL; };
  for (auto& I : inst) {
    mix((unsigned long long)I->S.nx[0]);
    mix((unsigned long long)I->S.nx[1]);
    mix((unsigned long long)I->S.blocks.size());
  }
  char key[256];
  snprintf(key, sizeof(key), "v2:%s:%s:s%d:n%zu:%016llx", prop.gcnArchName, sizeof(R) == 4 ? "fp32" : "fp64", sbp,
           inst.size(), h);
  std::string cache;
  if (const char* e = getenv("FDF_TUNE_CACHE")) cache = e;
  else if (const char* home = getenv("HOME")) cache = std::string(home) + "/.cache/fdfault_amd_tune.txt";
  int ftw = 0, fcx = 0, fp = -1;
  if (const char* e = getenv("FDF_BATCH_TW")) ftw = atoi(e);
  if (const char* e = getenv("FDF_BATCH_CX")) fcx = atoi(e);
  if (const char* e = getenv("FDF_BATCH_PERSIST")) fp = atoi(e) ? 1 : 0;
  if (ftw && fcx) {  // fixed by the caller
    build(ftw, fcx, fp < 0 ? 1 : fp);
    return;
  }
  if (!cache.empty() && fp < 0) {
    if (FILE* fh = fopen(cache.c_str(), "r")) {
      char line[512], k[256];
      int a = 0, c = 0, m = 0;
      double ms = 0.;
      while (fgets(line, sizeof(line), fh))
        if (sscanf(line, "%255s %d %d %d %lf", k, &a, &c, &m, &ms) == 5 && !strcmp(k, key) &&
            (a == 64 || a == 128 || a == 256) && c > 0 && (m == 0 || m == 1)) {
          ftw = a;
          fcx = c;
          fp = m;
          tune_ms = ms;
        }
      fclose(fh);
    }
  }
  if (ftw) {
    tuned_from_cache = 1;
    build(ftw, fcx, fp);
    return;
  }
  // save df of every instance, time the candidates, restore
  std::vector<DBuf<R>> saved(inst.size());
  for (size_t i = 0; i < inst.size(); i++) {
    saved[i].alloc(inst[i]->df.n);
    HIPCHK(hipMemcpyAsync(saved[i].p, inst[i]->df.p, inst[i]->df.n * sizeof(R), hipMemcpyDeviceToDevice, stream));
  }
  double best = 1e300;
  int btw = 256, bcx = 8, bp = 1;
  for (int m : {1, 0}) {
    if (fp >= 0 && m != fp) continue;
    for (int a : {64, 128, 256})
      for (int c : {1, 2, 4, 8, 16}) {
        const double ms = time_config(a, c, m);
        if (ms < best) {
          best = ms;
          btw = a;
          bcx = c;
          bp = m;
        }
      }
  }
  for (size_t i = 0; i < inst.size(); i++)
    HIPCHK(hipMemcpyAsync(inst[i]->df.p, saved[i].p, inst[i]->df.n * sizeof(R), hipMemcpyDeviceToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  tune_ms = best;
  build(btw, bcx, bp);
  if (!cache.empty() && fp < 0)
    if (FILE* fh = fopen(cache.c_str(), "a")) {
      fprintf(fh, "%s %d %d %d %.6f\n", key, btw, bcx, bp, best);
      fclose(fh);
    }
}

template <typename R>
void Batch<R>::step(int n) {
  for (int k = 0; k < n; k++) {
    if (opts.use_graph) {
      const int par = cur;
      if (!graphs[par]) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int st = 0; st < nst; st++) launch_stage(st);
        hipLaunchKernelGGL(advance_step, dim3(1), dim3(1), 0, stream, dstep.p);
        hipGraph_t g;
        HIPCHK(hipStreamEndCapture(stream, &g));
        HIPCHK(hipGraphInstantiate(&graphs[par], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
        cur = par;
      }
      if (!dstep_valid) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(dstep.p, &cur_step, sizeof(int), hipMemcpyHostToDevice));
        dstep_valid = true;
      }
      HIPCHK(hipGraphLaunch(graphs[par], stream));
      cur ^= (nst & 1);
    } else {
      if (!dstep_valid) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(hipMemcpy(dstep.p, &cur_step, sizeof(int), hipMemcpyHostToDevice));
        dstep_valid = true;
      }
      for (int st = 0; st < nst; st++) launch_stage(st);
      hipLaunchKernelGGL(advance_step, dim3(1), dim3(1), 0, stream, dstep.p);
    }
    cur_step++;
    const bool stale = !(stage_dfmode(nst - 1) & DF_STORE);
    for (auto& I : inst) I->df_stale = stale;
    if (opts.write_outputs)
      for (auto& I : inst) {  // gathered on the batch stream, after this step's launches
        I->cur = cur;
        I->cur_step = cur_step;
        I->write_outputs(cur_step, stream);
      }
  }
  sync();
}

// wait for the batch stream, check the error flags, and hand the current
// buffer parity and step to every instance (reads go through them)
template <typename R>
void Batch<R>::sync() {
  HIPCHK(hipStreamSynchronize(stream));
  int e = 0;
  HIPCHK(hipMemcpy(&e, err.p, sizeof(int), hipMemcpyDeviceToHost));
  for (auto& I : inst) {
    I->cur = cur;
    I->cur_step = cur_step;
  }
  if (e) {
    int z = 0;
    HIPCHK(hipMemcpy(err.p, &z, sizeof(int), hipMemcpyHostToDevice));
    throw Error{fault_error_message(e)};
  }
}

template <typename R>
void Batch<R>::finish() {
  sync();
  for (auto& I : inst) I->finish();
}

// ---- STZ learned proposal: model, log, counters --------------------------
template <typename R>
void Solver<R>::stz_refresh() {
  for (size_t i = 0; i < fdev.size(); i++) {
    if (S.ifaces[i].law != LAW_STZ) continue;
    fdev[i].mlp = stz_model.p;
    fdev[i].log = stz_log.p;
    fdev[i].log_n = stz_log.p ? stz_cnt.p : nullptr;
    fdev[i].log_cap = (long long)(stz_log.n / 5);
  }
  for (int q = 0; q < 2; q++)  // captured launches hold the old arguments
    if (graphs[q]) {
      HIPCHK(hipGraphExecDestroy(graphs[q]));
      graphs[q] = nullptr;
    }
}

template <typename R>
void Solver<R>::stz_set_model(const float* params, int n) {
  HIPCHK(hipStreamSynchronize(stream));
  if (!params) {
    stz_model.free();
  } else {
    if (n != MLP_NP) throw Error{"stz model: expected " + std::to_string(MLP_NP) + " parameters"};
    stz_model.upload(params, (size_t)n);
  }
  stz_refresh();
}

template <typename R>
void Solver<R>::stz_enable_log(long cap) {
  HIPCHK(hipStreamSynchronize(stream));
  if (!stz_cnt.p) throw Error{"stz log: no STZ interface"};
  if (cap <= 0) stz_log.free();
  else stz_log.alloc((size_t)cap * 5);
  const unsigned long long z = 0;
  HIPCHK(hipMemcpy(stz_cnt.p, &z, sizeof(z), hipMemcpyHostToDevice));
  stz_refresh();
}

template <typename R>
long Solver<R>::stz_read_log(double* out, long cap) {
  HIPCHK(hipStreamSynchronize(stream));
  if (!stz_cnt.p || !stz_log.p) return 0;
  unsigned long long n = 0;
  HIPCHK(hipMemcpy(&n, stz_cnt.p, sizeof(n), hipMemcpyDeviceToHost));
  long rows = (long)std::min<unsigned long long>(n, (unsigned long long)(stz_log.n / 5));
  if (out) {
    rows = std::min(rows, cap);
    HIPCHK(hipMemcpy(out, stz_log.p, (size_t)rows * 5 * sizeof(double), hipMemcpyDeviceToHost));
  }
  return rows;
}

// counters since the last call: solves, Newton iterations, proposals used,
// fallbacks to the plain solve
template <typename R>
void Solver<R>::stz_stats(unsigned long long out[4]) {
  HIPCHK(hipStreamSynchronize(stream));
  for (int q = 0; q < 4; q++) out[q] = 0;
  if (!stz_cnt.p) return;
  HIPCHK(hipMemcpy(out, stz_cnt.p + 1, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  const unsigned long long z[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpy(stz_cnt.p + 1, z, sizeof(z), hipMemcpyHostToDevice));
}

// ---- MLP training on the GPU ---------------------------------------------
// ---- Allocation-free gradient of the proposal network ----------------------
// Scratch of fdf_mlp_grad_noalloc, a file-scope device array each GPU holds
// until the process ends: per-sample activations and deltas (H1, H2, D1, D2:
// MLP_H floats each; DY: one) for up to MLP_NOALLOC_MAXN samples, the MLP_GS
// weight-gradient slices, and the loss sum.  Host code never passes the
// array's name to a kernel: it is a host-side stand-in, and the device
// address comes from hipGetSymbolAddress once per device.
constexpr long MLP_NOALLOC_MAXN = 4096;
constexpr size_t MLP_SCR_SAMPLE = 4 * MLP_H + 1;
constexpr size_t MLP_SCR_GP = (size_t)MLP_NOALLOC_MAXN * MLP_SCR_SAMPLE;
constexpr size_t MLP_SCR_LOSS = MLP_SCR_GP + (size_t)MLP_GS * MLP_NW;
constexpr size_t MLP_SCR_FLOATS = MLP_SCR_LOSS + 64;  // 4.36 MB
static __device__ __align__(256) float scratch[MLP_SCR_FLOATS];

static std::mutex g_scratch_mu;
static float* g_scratch_dev[64] = {};

// the scratch array's address on the current device (looked up on the
// first call there)
static float* scratch_address() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error{"fdf_mlp_grad_noalloc: device ordinal out of range"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error{std::string("fdf_mlp_grad_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e)};
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  return g_scratch_dev[dev];
}

__global__ void mlp_fill_kernel(float* __restrict__ p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void mlp_loss_mean_kernel(const float* __restrict__ sum, float* __restrict__ out, float n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = sum[0] / n;
}

struct MlpBufs {
  DBuf<float> P, G, GP, M, V, X, T, H1, H2, D1, D2, DY, loss;
  DBuf<int> idx;
};

static void mlp_alloc(MlpBufs& b, long n, int B) {
  b.G.alloc(MLP_NW);
  b.GP.alloc((size_t)MLP_GS * MLP_NW);
  b.H1.alloc((size_t)B * MLP_H);
  b.H2.alloc((size_t)B * MLP_H);
  b.D1.alloc((size_t)B * MLP_H);
  b.D2.alloc((size_t)B * MLP_H);
  b.DY.alloc((size_t)B);
  b.loss.alloc(1);
  (void)n;
}

// gradient of the batch loss (mean of (y-t)^2/2) for B rows of X/T selected
// by idx (or rows 0..B-1); the loss sum is added to b.loss
static void mlp_batch_grad(MlpBufs& b, const int* idx, int B, hipStream_t st) {
  hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3((B + 255) / 256), dim3(256), 0, st, b.P.p, b.X.p, b.T.p, idx, B, b.H1.p,
                     b.H2.p, b.D1.p, b.D2.p, b.DY.p, b.loss.p);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(2 * MLP_H + 1, MLP_GS), dim3(256), 0, st, b.X.p, idx, B, b.H1.p, b.H2.p,
                     b.D1.p, b.D2.p, b.DY.p, b.GP.p);
  hipLaunchKernelGGL(mlp_gsum_kernel, dim3((MLP_NW + 255) / 256), dim3(256), 0, st, b.GP.p, b.G.p);
  HIPCHK(hipGetLastError());
}

// Train the proposal network on log rows (phi, sn, chi, eta, log V): inputs
// and target standardised with the data's mean and deviation, Glorot-uniform
// initial weights, Adam over shuffled minibatches with the learning rate
// decaying geometrically to lr/10 over the run.  Returns the parameter
// vector (with normalisation) and the mean loss of each epoch.
static void mlp_train(const double* rows_in, long n, int epochs, int B, float lr, unsigned seed, float* params,
                      float* epoch_loss) {
  if (n < 1 || epochs < 1 || B < 1) throw Error{"mlp_train: empty data or bad settings"};
  if (B > n) B = (int)n;
  // the fault kernel logs its rows in atomic order; training sees them in
  // lexicographic order, so the statistics, the shuffled minibatches and the
  // trained network depend on the set of rows only
  std::vector<long> ord((size_t)n);
  for (long r = 0; r < n; r++) ord[r] = r;
  std::sort(ord.begin(), ord.end(), [&](long a, long b) {
    return std::lexicographical_compare(rows_in + a * 5, rows_in + a * 5 + 5, rows_in + b * 5, rows_in + b * 5 + 5);
  });
  std::vector<double> sorted_rows((size_t)n * 5);
  for (long r = 0; r < n; r++)
    for (int c = 0; c < 5; c++) sorted_rows[(size_t)r * 5 + c] = rows_in[ord[r] * 5 + c];
  const double* rows = sorted_rows.data();
  std::vector<float> P(MLP_NP, 0.f);
  double mean[5] = {0, 0, 0, 0, 0}, sd[5] = {0, 0, 0, 0, 0};
  for (long r = 0; r < n; r++)
    for (int c = 0; c < 5; c++) mean[c] += rows[r * 5 + c];
  for (int c = 0; c < 5; c++) mean[c] /= (double)n;
  for (long r = 0; r < n; r++)
    for (int c = 0; c < 5; c++) sd[c] += (rows[r * 5 + c] - mean[c]) * (rows[r * 5 + c] - mean[c]);
  for (int c = 0; c < 5; c++) {
    sd[c] = sqrt(sd[c] / (double)n);
    if (!(sd[c] > 1e-12 * (fabs(mean[c]) + 1e-300))) sd[c] = 1.;  // constant column
  }
  for (int c = 0; c < MLP_IN; c++) {
    P[MLP_NORM + c] = (float)mean[c];
    P[MLP_NORM + MLP_IN + c] = (float)sd[c];
  }
  P[MLP_NORM + 2 * MLP_IN] = (float)mean[4];
  P[MLP_NORM + 2 * MLP_IN + 1] = (float)sd[4];
  std::vector<float> X((size_t)n * MLP_IN), T((size_t)n);
  for (long r = 0; r < n; r++) {
    for (int c = 0; c < MLP_IN; c++) X[r * MLP_IN + c] = (float)((rows[r * 5 + c] - mean[c]) / sd[c]);
    T[r] = (float)((rows[r * 5 + 4] - mean[4]) / sd[4]);
  }
  std::mt19937 rng(seed);
  auto glorot = [&](int off, int fan_in, int fan_out, int count) {
    const float a = sqrtf(6.f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < count; i++) P[off + i] = u(rng);
  };
  glorot(MLP_W1, MLP_IN, MLP_H, MLP_H * MLP_IN);
  glorot(MLP_W2, MLP_H, MLP_H, MLP_H * MLP_H);
  glorot(MLP_W3, MLP_H, 1, MLP_H);
  MlpBufs b;
  b.P.upload(P);
  b.X.upload(X);
  b.T.upload(T);
  mlp_alloc(b, n, B);
  b.M.alloc(MLP_NW);
  b.V.alloc(MLP_NW);
  hipStream_t st;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  b.M.zero(st);
  b.V.zero(st);
  std::vector<int> perm((size_t)n);
  for (long r = 0; r < n; r++) perm[r] = (int)r;
  DBuf<int> didx;
  didx.alloc((size_t)n);
  const long nb = n / B;  // full batches per epoch
  int stepc = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    HIPCHK(hipMemcpyAsync(didx.p, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(b.loss.p, 0, sizeof(float), st));
    const float lre = lr * powf(0.1f, (float)ep / (float)epochs);
    for (long k = 0; k < nb; k++) {
      mlp_batch_grad(b, didx.p + k * B, B, st);
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NW + 255) / 256), dim3(256), 0, st, b.P.p, b.G.p, b.M.p, b.V.p,
                         lre, 0.9f, 0.999f, 1e-8f, ++stepc);
    }
    HIPCHK(hipGetLastError());
    float ls = 0.f;
    HIPCHK(hipMemcpyAsync(&ls, b.loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (epoch_loss) epoch_loss[ep] = ls / (float)(nb * B);
  }
  HIPCHK(hipMemcpyAsync(P.data(), b.P.p, MLP_NW * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipStreamDestroy(st));
  std::copy(P.begin(), P.end(), params);
}

}  // namespace fdf

struct fdf_s {
  std::unique_ptr<fdf::SolverBase> solver;
};

struct fdf_batch_s {
  std::unique_ptr<fdf::BatchBase> b;
};

using fdf::Error;

template <typename Fn>
static int guarded(Fn&& fn) {
  try {
    fn();
    return 0;
  } catch (Error& e) {
    fdf::g_last_error = e.msg;
  } catch (std::exception& e) {
    fdf::g_last_error = e.what();
  } catch (...) {
    fdf::g_last_error = "unknown error";
  }
  return -1;
}

// Entry points turn a null handle or a null buffer into an error return
// rather than a crash.
static fdf::SolverBase* solver_of(fdf_t* h, const char* fn) {
  if (!h || !h->solver) throw Error{std::string(fn) + ": null handle"};
  return h->solver.get();
}

static fdf::BatchBase* batch_of(fdf_batch_t* h, const char* fn) {
  if (!h || !h->b) throw Error{std::string(fn) + ": null handle"};
  return h->b.get();
}

static void need(const void* p, const char* fn) {
  if (!p) throw Error{std::string(fn) + ": null argument"};
}

extern "C" {

const char* fdf_last_error(void) { return fdf::g_last_error.c_str(); }

int fdf_open(const char* in_path, const fdf_opts* opts, fdf_t** out) {
  if (!in_path || !out) {
    fdf::g_last_error = "fdf_open: null argument";
    return -2;
  }
  *out = nullptr;
  fdf_opts o{};
  o.quiet = 1;
  if (opts) o = *opts;
  return guarded([&] {
    std::unique_ptr<fdf_s> h(new fdf_s());
    if (o.precision == FDF_FP32) h->solver.reset(new fdf::Solver<float>(in_path, o));
    else if (o.precision == FDF_FP64) h->solver.reset(new fdf::Solver<double>(in_path, o));
    else throw Error{"fdf_open: unknown precision"};
    *out = h.release();
  });
}

void fdf_close(fdf_t* h) { delete h; }

int fdf_step(fdf_t* h, int nsteps) {
  return guarded([&] { solver_of(h, "fdf_step")->step(nsteps); });
}

int fdf_stage(fdf_t* h, int istep, int stage) {
  return guarded([&] {
    fdf::SolverBase* sv = solver_of(h, "fdf_stage");
    if (stage < 0 || stage >= sv->S.rk.nstages) throw Error{"fdf_stage: bad stage"};
    sv->stage(istep, stage);
  });
}

int fdf_run(const char* in_path, const fdf_opts* opts) {
  fdf_opts o{};
  if (opts) o = *opts;
  o.write_outputs = 1;
  fdf_t* h = nullptr;
  int rc = fdf_open(in_path, &o, &h);
  if (rc) return rc;
  rc = fdf_step(h, h->solver->S.nt);
  if (!rc) rc = fdf_finish(h);
  fdf_close(h);
  return rc;
}

int fdf_read_field(fdf_t* h, int field, const int lo[3], const int hi[3], const int st[3], double* host) {
  return guarded([&] {
    fdf::SolverBase* sv = solver_of(h, "fdf_read_field");
    for (const void* p : {(const void*)lo, (const void*)hi, (const void*)st, (const void*)host}) need(p, "fdf_read_field");
    sv->read_field(field, lo, hi, st, host);
  });
}

int fdf_read_array(fdf_t* h, int which, double* host, long n) {
  return guarded([&] {
    need(host, "fdf_read_array");
    solver_of(h, "fdf_read_array")->read_array(which, host, n);
  });
}

int fdf_write_array(fdf_t* h, int which, const double* host, long n) {
  return guarded([&] {
    need(host, "fdf_write_array");
    solver_of(h, "fdf_write_array")->write_array(which, host, n);
  });
}

int fdf_save_state(fdf_t* h, const char* path) {
  return guarded([&] {
    need(path, "fdf_save_state");
    solver_of(h, "fdf_save_state")->save_state(path);
  });
}

int fdf_load_state(fdf_t* h, const char* path) {
  return guarded([&] {
    need(path, "fdf_load_state");
    solver_of(h, "fdf_load_state")->load_state(path);
  });
}

long fdf_read_iface(fdf_t* h, int iface, int which, double* host, long n) {
  long r = -1;
  int rc = guarded([&] { r = solver_of(h, "fdf_read_iface")->read_iface(iface, which, host, n); });
  return rc ? -1 : r;
}

long fdf_read_front(fdf_t* h, int iface, double* host, long n) {
  long r = -1;
  int rc = guarded([&] { r = solver_of(h, "fdf_read_front")->read_front(iface, host, n); });
  return rc ? -1 : r;
}

int fdf_finish(fdf_t* h) {
  return guarded([&] { solver_of(h, "fdf_finish")->finish(); });
}

static void fill_info(const fdf::SolverBase* sv, fdf_info* info) {
  const fdf::SetupH& S = sv->S;
  info->nt = S.nt;
  info->nstages = S.rk.nstages;
  info->ndim = S.ndim;
  info->mode = S.mode;
  for (int i = 0; i < 3; i++) info->nx[i] = S.nx[i];
  info->nfields = S.nf;
  info->nfields_total = S.nf + S.nfp;
  info->nifaces = (int)S.ifaces.size();
  info->sbporder = S.sbporder;
  info->plastic = S.plastic;
  info->nblocks = (int)S.blocks.size();
  info->step = sv->cur_step;
  info->precision = sv->opts.precision;
  info->dt = S.dt;
  info->ttot = S.ttot;
  info->cfl = S.cfl;
  info->t = sv->cur_step * S.dt;
  info->tiled_blocks = sv->ntiled;
  info->row_blocks = sv->nrows;
  info->curv_blocks = sv->ncurv;
  info->wide_blocks = sv->nwide;
  info->fused_faults = sv->nfused;
}

int fdf_get_info(fdf_t* h, fdf_info* info) {
  if (!h || !info) return -2;
  fill_info(h->solver.get(), info);
  return 0;
}

int fdf_open_batch(const char* const* paths, int n, const fdf_opts* opts, fdf_batch_t** out) {
  if (!paths || n <= 0 || !out) {
    fdf::g_last_error = "fdf_open_batch: null argument or empty batch";
    return -2;
  }
  *out = nullptr;
  fdf_opts o{};
  o.quiet = 1;
  if (opts) o = *opts;
  return guarded([&] {
    std::vector<std::string> v;
    for (int i = 0; i < n; i++) {
      if (!paths[i]) throw Error{"fdf_open_batch: null path"};
      v.push_back(paths[i]);
    }
    std::unique_ptr<fdf_batch_s> h(new fdf_batch_s());
    if (o.precision == FDF_FP32) h->b.reset(new fdf::Batch<float>(v, o));
    else if (o.precision == FDF_FP64) h->b.reset(new fdf::Batch<double>(v, o));
    else throw Error{"fdf_open_batch: unknown precision"};
    *out = h.release();
  });
}

void fdf_batch_close(fdf_batch_t* h) { delete h; }

int fdf_batch_size(fdf_batch_t* h) { return h ? h->b->size() : -2; }

int fdf_batch_step(fdf_batch_t* h, int nsteps) {
  return guarded([&] { batch_of(h, "fdf_batch_step")->step(nsteps); });
}

int fdf_batch_get_info(fdf_batch_t* h, int inst, fdf_info* info) {
  if (!h || !info || inst < 0 || inst >= h->b->size()) return -2;
  fill_info(h->b->instance(inst), info);
  info->fused_faults = h->b->fused_instances() ? 1 : 0;  // the batch's own launches
  return 0;
}

long fdf_batch_read_iface(fdf_batch_t* h, int inst, int iface, int which, double* host, long n) {
  long r = -1;
  int rc = guarded([&] {
    fdf::BatchBase* b = batch_of(h, "fdf_batch_read_iface");
    if (inst < 0 || inst >= b->size()) throw Error{"fdf_batch_read_iface: bad instance"};
    r = b->instance(inst)->read_iface(iface, which, host, n);
  });
  return rc ? -1 : r;
}

long fdf_batch_read_front(fdf_batch_t* h, int inst, int iface, double* host, long n) {
  long r = -1;
  int rc = guarded([&] {
    fdf::BatchBase* b = batch_of(h, "fdf_batch_read_front");
    if (inst < 0 || inst >= b->size()) throw Error{"fdf_batch_read_front: bad instance"};
    r = b->instance(inst)->read_front(iface, host, n);
  });
  return rc ? -1 : r;
}

int fdf_batch_read_array(fdf_batch_t* h, int inst, int which, double* host, long n) {
  return guarded([&] {
    fdf::BatchBase* b = batch_of(h, "fdf_batch_read_array");
    need(host, "fdf_batch_read_array");
    if (inst < 0 || inst >= b->size()) throw Error{"fdf_batch_read_array: bad instance"};
    b->instance(inst)->read_array(which, host, n);
  });
}

int fdf_batch_tuning(fdf_batch_t* h, int* tile_width, int* chunk, double* stage_ms, int* from_cache) {
  if (!h) return -2;
  if (tile_width) *tile_width = h->b->tw;
  if (chunk) *chunk = h->b->cx;
  if (stage_ms) *stage_ms = h->b->tune_ms;
  if (from_cache) *from_cache = h->b->tuned_from_cache;
  return 0;
}

int fdf_batch_persistent(fdf_batch_t* h) { return h ? h->b->persist : -2; }

int fdf_batch_finish(fdf_batch_t* h) {
  return guarded([&] { batch_of(h, "fdf_batch_finish")->finish(); });
}

int fdf_mlp_grad(const float* params, const float* X, const float* T, long n, float* grad, float* loss) {
  return guarded([&] {
    if (!params || !X || !T || n < 1 || n > (1L << 24)) throw Error{"fdf_mlp_grad: bad arguments"};
    fdf::MlpBufs b;
    b.P.upload(params, fdf::MLP_NP);
    b.X.upload(X, (size_t)n * fdf::MLP_IN);
    b.T.upload(T, (size_t)n);
    fdf::mlp_alloc(b, n, (int)n);
    HIPCHK(hipMemset(b.loss.p, 0, sizeof(float)));
    fdf::mlp_batch_grad(b, nullptr, (int)n, nullptr);
    HIPCHK(hipDeviceSynchronize());
    if (grad) HIPCHK(hipMemcpy(grad, b.G.p, fdf::MLP_NW * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) {
      float l = 0.f;
      HIPCHK(hipMemcpy(&l, b.loss.p, sizeof(float), hipMemcpyDeviceToHost));
      *loss = l / (float)n;
    }
  });
}

int fdf_mlp_grad_noalloc(int device, const float* params, const float* X, const float* T, long n, float* grad,
                         float* loss, void* stream) {
  return guarded([&] {
    if (!params || !X || !T || !grad) throw Error{"fdf_mlp_grad_noalloc: null argument"};
    if (n < 1 || n > fdf::MLP_NOALLOC_MAXN)
      throw Error{"fdf_mlp_grad_noalloc: batch of " + std::to_string(n) + " samples outside 1.." +
                  std::to_string(fdf::MLP_NOALLOC_MAXN) + " (the fixed scratch holds at most that many)"};
    HIPCHK(hipSetDevice(device));
    float* scr = fdf::scratch_address();
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int B = (int)n;
    float* H1 = scr;
    float* H2 = H1 + (size_t)B * fdf::MLP_H;
    float* D1 = H2 + (size_t)B * fdf::MLP_H;
    float* D2 = D1 + (size_t)B * fdf::MLP_H;
    float* DY = D2 + (size_t)B * fdf::MLP_H;
    float* GP = scr + fdf::MLP_SCR_GP;
    float* lsum = scr + fdf::MLP_SCR_LOSS;
    // the loss sum starts at zero (a fill kernel, not a memset, so that the
    // sequence can be captured in a graph)
    hipLaunchKernelGGL(fdf::mlp_fill_kernel, dim3(1), dim3(64), 0, st, lsum, 1, 0.f);
    hipLaunchKernelGGL(fdf::mlp_fwd_bwd_kernel, dim3((B + 255) / 256), dim3(256), 0, st, params, X, T, nullptr, B, H1,
                       H2, D1, D2, DY, lsum);
    hipLaunchKernelGGL(fdf::mlp_grad_kernel, dim3(2 * fdf::MLP_H + 1, fdf::MLP_GS), dim3(256), 0, st, X, nullptr, B,
                       H1, H2, D1, D2, DY, GP);
    hipLaunchKernelGGL(fdf::mlp_gsum_kernel, dim3((fdf::MLP_NW + 255) / 256), dim3(256), 0, st, GP, grad);
    if (loss) hipLaunchKernelGGL(fdf::mlp_loss_mean_kernel, dim3(1), dim3(64), 0, st, lsum, loss, (float)n);
    HIPCHK(hipGetLastError());
  });
}

int fdf_mlp_predict(const float* params, const float* X, long n, float* Y) {
  return guarded([&] {
    if (!params || !X || !Y || n < 1) throw Error{"fdf_mlp_predict: bad arguments"};
    fdf::DBuf<float> P, dx, dy;
    P.upload(params, fdf::MLP_NP);
    dx.upload(X, (size_t)n * fdf::MLP_IN);
    dy.alloc((size_t)n);
    hipLaunchKernelGGL(fdf::mlp_predict_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, P.p, dx.p,
                       (int)n, dy.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(Y, dy.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int fdf_mlp_predict_mfma(const float* params, const float* X, long n, float* Y) {
  return guarded([&] {
    if (!params || !X || !Y || n < 1) throw Error{"fdf_mlp_predict_mfma: bad arguments"};
    fdf::DBuf<float> P, dx, dy;
    P.upload(params, fdf::MLP_NP);
    dx.upload(X, (size_t)n * fdf::MLP_IN);
    dy.alloc((size_t)n);
    hipLaunchKernelGGL(fdf::mlp_predict_wave_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, P.p,
                       dx.p, (int)n, dy.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(Y, dy.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int fdf_mlp_train(const double* rows, long n, int epochs, int batch, float lr, unsigned seed, float* params,
                  float* epoch_loss) {
  return guarded([&] {
    if (!rows || !params) throw Error{"fdf_mlp_train: null argument"};
    fdf::mlp_train(rows, n, epochs, batch, lr, seed, params, epoch_loss);
  });
}

int fdf_mlp_param_count(void) { return fdf::MLP_NP; }

int fdf_stz_set_model(fdf_t* h, const float* params, int n) {
  return guarded([&] {
    if (auto* s = dynamic_cast<fdf::Solver<double>*>(solver_of(h, "fdf_stz"))) s->stz_set_model(params, n);
    else if (auto* s32 = dynamic_cast<fdf::Solver<float>*>(solver_of(h, "fdf_stz"))) s32->stz_set_model(params, n);
  });
}

int fdf_stz_log(fdf_t* h, long capacity) {
  return guarded([&] {
    if (auto* s = dynamic_cast<fdf::Solver<double>*>(solver_of(h, "fdf_stz"))) s->stz_enable_log(capacity);
    else if (auto* s32 = dynamic_cast<fdf::Solver<float>*>(solver_of(h, "fdf_stz"))) s32->stz_enable_log(capacity);
  });
}

long fdf_stz_log_read(fdf_t* h, double* out, long cap) {
  long r = -1;
  int rc = guarded([&] {
    if (auto* s = dynamic_cast<fdf::Solver<double>*>(solver_of(h, "fdf_stz"))) r = s->stz_read_log(out, cap);
    else if (auto* s32 = dynamic_cast<fdf::Solver<float>*>(solver_of(h, "fdf_stz"))) r = s32->stz_read_log(out, cap);
  });
  return rc ? -1 : r;
}

int fdf_stz_stats(fdf_t* h, unsigned long long* out4) {
  return guarded([&] {
    if (!out4) throw Error{"fdf_stz_stats: null argument"};
    if (auto* s = dynamic_cast<fdf::Solver<double>*>(solver_of(h, "fdf_stz"))) s->stz_stats(out4);
    else if (auto* s32 = dynamic_cast<fdf::Solver<float>*>(solver_of(h, "fdf_stz"))) s32->stz_stats(out4);
  });
}


double fdf_bulk_kernel_ms(fdf_t* h, int* nlaunch) { return h && h->solver ? h->solver->bulk_ms(nlaunch) : -1.; }
double fdf_fault_kernel_ms(fdf_t* h, int* nstages) { return h && h->solver ? h->solver->fault_ms(nstages) : -1.; }

int fdf_host_setup(const char* in_path, fdf_info* info, double* x, double* metric, double* jac) {
  return guarded([&] {
    fdf::SetupH S;
    fdf::compile_problem(in_path, S, false, true);
    if (info) {
      memset(info, 0, sizeof(*info));
      info->nt = S.nt;
      info->nstages = S.rk.nstages;
      info->ndim = S.ndim;
      info->mode = S.mode;
      for (int i = 0; i < 3; i++) info->nx[i] = S.nx[i];
      info->nfields = S.nf;
      info->nfields_total = S.nf + S.nfp;
      info->nifaces = (int)S.ifaces.size();
      info->sbporder = S.sbporder;
      info->plastic = S.plastic;
      info->nblocks = (int)S.blocks.size();
      info->dt = S.dt;
      info->ttot = S.ttot;
      info->cfl = S.cfl;
    }
    if (x) memcpy(x, S.x.data(), S.x.size() * sizeof(double));
    if (metric) memcpy(metric, S.metric.data(), S.metric.size() * sizeof(double));
    if (jac) memcpy(jac, S.jac.data(), S.jac.size() * sizeof(double));
  });
}

}  // extern "C"
