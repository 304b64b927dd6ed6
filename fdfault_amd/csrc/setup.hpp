// Host-side problem compiler of the MI355X time-stepping core.
//
// Reads a reference .in file and everything it names (surfaces, load,
// state, parameter, heterogeneous stress and material files), and builds the
// flat, device-ready description of the problem: global SoA grid arrays
// (x, metric, jac), per-block stencil ranges and material constants,
// per-face boundary normals, per-interface normals, impedances, friction
// parameters and perturbation shape values.  Everything here runs once, on
// the host, in FP64; the arithmetic of each derived quantity follows the
// reference routine cited at its definition so the device starts from
// bit-identical inputs.
#pragma once

#include <cstdint>
#include <new>
#include <string>
#include <utility>
#include <unordered_map>
#include <vector>

namespace fdf {

// Allocator whose value-initialisation leaves doubles uninitialised: the
// grid arrays (gigabytes on large grids) are first written by the parallel
// set-up loops, which then also place their pages, instead of being zeroed
// by one thread on resize.
template <typename T>
struct uninit_alloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = uninit_alloc<U>;
  };
  uninit_alloc() = default;
  template <typename U>
  uninit_alloc(const uninit_alloc<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
using gvec = std::vector<double, uninit_alloc<double>>;

struct Error {
  std::string msg;
};

// SBP first-derivative and dissipation tables (fd.cpp:33-155); row 0 is
// the interior stencil, rows 1..2s-2 the one-sided boundary rows.
struct SBP {
  int s = 0;
  double h0 = 0.;
  double c[7][9] = {};
  double d[7][9] = {};
};

struct LowStorageRK {  // rk.cpp:8-98
  int nstages = 0;
  double A[5] = {}, B[5] = {}, C[6] = {};
};

struct FaceBC {  // one external boundary face of a block (boundary.cpp:15-240)
  bool active = false;
  double r = 0., alpha = 0., theta = 0.;
  double cp = 0., cs = 0., zp = 0., zs = 0., gamma = 0.;
  int n0 = 0, n1 = 0;
  std::vector<double> nrm;  // [ndim][n0][n1], outward unit normal
  std::vector<double> dl;   // [n0][n1], |grad xi| / (h0 * dxi)
};

struct BlockH {
  int coords[3] = {0, 0, 0};
  int nx[3] = {1, 1, 1}, xm[3] = {0, 0, 0};
  int mlb[3], mc[3], mrb[3], prb[3];
  double dx[3];
  double rho = 0., lam = 0., G = 0.;
  double pmu = 0.5735, pc = 0., pbeta = 0.2867, peta = 0.2775;  // material.cpp:15-21
  double cp = 0., cs = 0., zp = 0., zs = 0.;
  double x0[3] = {0., 0., 0.}, L[3] = {0., 0., 0.};
  std::string btype[6], bfile[6];
  FaceBC face[6];
  double K[3][4];  // per direction: dt/rho/dx, dt*(2G+lam)/dx, dt*G/dx, dt*lam/dx (set with dt)
};

// a spatial perturbation (pert.cpp:7-148); xy = shape value per fault point
struct PertH {
  double t0 = 0.;
  double amp[9] = {};
  std::vector<double> xy;
};

enum { LAW_LOCKED = 0, LAW_FRICTIONLESS = 1, LAW_SLIPWEAK = 2, LAW_STZ = 3 };

struct IfaceH {
  int law = 0, dir = 0;
  int b1 = 0, b2 = 0;  // linear block indices
  int n0 = 0, n1 = 0;
  int org[3] = {0, 0, 0};    // global index of the first side-1 point
  int xm[3], xp[3];          // interface box in global indices (block-2 side)
  double cp1, cs1, zp1, zs1, cp2, cs2, zp2, zs2, gamma1, gamma2;
  std::vector<double> nrm;   // [ndim][n0][n1] unit normal (side-1 metric)
  std::vector<double> dl1, dl2;
  std::vector<PertH> loads;  // amplitudes sn s2 s3
  bool load_file = false;
  std::vector<double> lf[3];
  std::vector<PertH> perts;  // slipweak: dc mus mud c0 trup tc ; stz: v0 f0 a muy c0 R beta chiw v1
  bool param_file = false;
  std::vector<double> pf[9];
  std::vector<double> state0;  // STZ initial effective temperature
};

struct OutputH {
  std::string name, field;
  int fidx = 0;       // field index into f, or interface array id
  int iface = -1;     // -1 grid field
  int comp = 0;       // component of interface vector arrays
  int tm = 0, tp = 0, ts = 1;
  int xm[3], xp[3], xs[3], nx[3];
};

struct SetupH;

// Device-side grid set-up, implemented by the runtime: the grid arrays (x,
// metric, Jacobian) of every block are built on the GPU and stay there; the
// host keeps the values of the points it needs (faces, interfaces, outputs),
// gathered on request.
struct GridBuilder {
  virtual ~GridBuilder() {}
  // transfinite interpolation of block B from its 2 ndim face surfaces
  // ([ndim][n0][n1] each) and its metric and Jacobian by the SBP operator
  virtual void build_block(const SetupH& S, const BlockH& B, const std::vector<double>* surf, const int* n0,
                           const int* n1) = 0;
  // false if any Jacobian of the blocks built so far is not positive
  virtual bool jacobians_positive() = 0;
  // x (ndim per point), metric (ndim^2) and Jacobian at global points idx
  virtual void gather(const SetupH& S, const std::vector<size_t>& idx, double* x, double* metric, double* jac) = 0;
  // max over block B's points and directions l of cs |grad xi_l| / dxi_l
  virtual double block_worst(const SetupH& S, const BlockH& B) = 0;
};

struct SetupH {
  std::string name, datadir;
  int nt = 0, ninfo = 1, rkorder = 0;
  double dt = 0., ttot = 0., cfl = 0.;
  LowStorageRK rk;
  SBP fd;
  int ndim = 0, mode = 0, nx[3] = {1, 1, 1}, nb[3] = {1, 1, 1};
  int sbporder = 0;
  bool plastic = false, plastic_tensor = false;
  double s0[6] = {0., 0., 0., 0., 0., 0.};
  int nf = 0, nfp = 0, ns = 0, nv = 0, nmat = 0, sidx[6] = {0, 1, 2, 3, 4, 5};
  bool hetstress = false, hetmat = false;
  double cdiss = 0.;
  size_t N = 0;
  gvec x, metric, jac;  // host set-up: every entry is written by the block set-up
  // device set-up: the grid arrays live on the GPU; gx/gm/gj hold the values
  // of the gathered points, slot gslot[g] of global point g
  GridBuilder* gb = nullptr;
  std::unordered_map<size_t, size_t> gslot;
  std::vector<double> gx, gm, gj;
  void gather(const std::vector<size_t>& idx);  // adds the points not gathered yet
  double x_at(int c, size_t g) const {
    return gb ? gx[gslot.at(g) * ndim + c] : x[(size_t)c * N + g];
  }
  double metric_at(int lm, size_t g) const {
    return gb ? gm[gslot.at(g) * ndim * ndim + lm] : metric[(size_t)lm * N + g];
  }
  double jac_at(size_t g) const { return gb ? gj[gslot.at(g)] : jac[g]; }
  std::vector<double> s, mat;
  std::vector<BlockH> blocks;
  std::vector<IfaceH> ifaces;
  std::vector<OutputH> outputs;
  bool has_front = false;
  double front_value = 0.;
  int front_field = 0;  // 0 V, 1 U (U never records: front.cpp:116-120)

  size_t idx(int i, int j, int k) const { return ((size_t)i * nx[1] + j) * nx[2] + k; }
};

// Parse + build; throws Error.  `outputs` parses the output list (else
// skipped).  With gb the grid arrays are built on the device (GridBuilder).
void compile_problem(const std::string& path, SetupH& S, bool outputs, bool quiet, GridBuilder* gb = nullptr);

// Whether block B's metric is diagonal and uniform and its Jacobian uniform,
// to 1e-9 relative (the SBP derivative of TFI coordinates leaves ~1e-12
// rounding noise); mdiag receives the diagonal at the block's centre point.
bool block_metric_uniform(const SetupH& S, const BlockH& B, double mdiag[3]);

}  // namespace fdf
