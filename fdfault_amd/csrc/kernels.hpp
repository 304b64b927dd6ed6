// HIP kernels of the MI355X time-stepping core (gfx950, wave64).
//
// One RK stage is two launches per interface / block:
//   fault_kernel : one thread per fault point; gathers both sides of the
//                  interface from f (read-only this stage), rotates into the
//                  fault frame, solves the friction law in FP64, writes the
//                  SAT terms for both sides into a small per-interface buffer
//                  and advances slip/state with the low-storage RK update.
//   bulk_kernel  : one thread per grid point; df = A*df + dt*RHS(f) with the
//                  SBP interior/boundary rows, then the external-boundary SAT
//                  of the point's faces, then the interface SAT terms written
//                  by fault_kernel, then f_new = f + B*df into the other f
//                  buffer (ping-pong), so f and df are each read and written
//                  once per stage.
// In FP64 every expression keeps the association of the reference routine it
// restates (cited at each function) and the code is compiled with
// -ffp-contract=off, so results track the reference to the last bit.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mlp.hpp"

namespace fdf {

// SBP tables in constant memory: row 0 interior, rows 1.. boundary rows
__constant__ double c_fd[7][9];
__constant__ double c_diss[7][9];

struct FaceDev {  // external boundary face (boundary.cpp:15-240)
  int active;
  int n0, n1;
  double r, alpha, theta, cp, cs, zp, zs, gamma;
  const double* nrm;  // [ndim][n0][n1] outward unit normal
  const double* dl;   // [n0][n1]
};

struct IfFaceDev {    // an interface face touching this block
  int side;           // 0: this block is block 1 (minus side), 1: block 2
  int dir;            // interface normal direction
  int fixed;          // global index along dir of the face row in this block
  int a, b;           // in-plane axes
  int org_a, org_b;   // global in-plane origin
  int n;              // points in the interface
  int n1;
  const double* sat;  // [2 sides][2 terms][nf][n]
};

struct BlockDev {
  int lo[3], n[3];    // first point and extents (global indices)
  int mc[3], mrb[3];  // end of left boundary rows, start of right boundary rows
  int hi[3];          // lo + n
  double K[3][4];     // per direction: dt/rho/dx, dt(2G+lam)/dx, dt G/dx, dt lam/dx
  double dx[3];
  double nu;          // plane-strain factor lam/(2(lam+G)) for plastic mode II
  double lam, G;
  double pmu, pc, pbeta, peta;
  int nif;
  FaceDev face[6];
  IfFaceDev ifc[6];
};

struct GridDev {
  int ndim, mode, nf, s;
  int Ny, Nz;
  long long N;
  int plastic, hetmat, hetstress;
  double cdiss;
  double s0[6];
  int ns, nv, sidx[6];
  int nfp, ptensor;
};

// Sub-boxes of a block a bulk_kernel launch covers (the whole block, or the
// shell around the tiled interior): thread t maps into them by prefix offsets.
struct Boxes {
  int nbox;
  int lo[6][3], n[6][3];
  long long start[7];
};

// Which of the stage's df/f traffic a bulk launch moves (bits): DF_LOAD reads
// df and scales it by A (fields::scale_df, fields.cpp:293-300; skipped where
// A = 0, the first stage of every low-storage scheme of rk.cpp:42-91),
// DF_STORE writes the new df (skipped on the last stage of a full time step,
// whose df only the next step's first stage would read, times A = 0), F_STORE
// writes the updated f (off when a skipped df is recomputed for a reader).
enum { DF_LOAD = 1, DF_STORE = 2, F_STORE = 4, DF_ALL = 7 };

struct StageDev {
  double A, B, dt;
  int roundtrip;  // end of an elastic step: the set_stress/remove_stress rounding of problem.cpp:160-164
  int dfmode = DF_ALL;
};

// rotated boundary fields, boundary.cpp:505-568
struct BFd {
  double v1, v2, v3, s11, s12, s13, s22, s23, s33;
};

__device__ __forceinline__ BFd rot_to_nt(const BFd& b, const double* nn, const double* t1, const double* t2) {
  BFd o;
  o.v1 = nn[0] * b.v1 + nn[1] * b.v2 + nn[2] * b.v3;
  o.v2 = t1[0] * b.v1 + t1[1] * b.v2 + t1[2] * b.v3;
  o.v3 = t2[0] * b.v1 + t2[1] * b.v2 + t2[2] * b.v3;
  o.s11 = (nn[0] * (b.s11 * nn[0] + b.s12 * nn[1] + b.s13 * nn[2]) + nn[1] * (b.s12 * nn[0] + b.s22 * nn[1] + b.s23 * nn[2]) +
           nn[2] * (b.s13 * nn[0] + b.s23 * nn[1] + b.s33 * nn[2]));
  o.s12 = (nn[0] * (b.s11 * t1[0] + b.s12 * t1[1] + b.s13 * t1[2]) + nn[1] * (b.s12 * t1[0] + b.s22 * t1[1] + b.s23 * t1[2]) +
           nn[2] * (b.s13 * t1[0] + b.s23 * t1[1] + b.s33 * t1[2]));
  o.s13 = (nn[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) + nn[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           nn[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  o.s22 = (t1[0] * (b.s11 * t1[0] + b.s12 * t1[1] + b.s13 * t1[2]) + t1[1] * (b.s12 * t1[0] + b.s22 * t1[1] + b.s23 * t1[2]) +
           t1[2] * (b.s13 * t1[0] + b.s23 * t1[1] + b.s33 * t1[2]));
  o.s23 = (t1[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) + t1[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           t1[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  o.s33 = (t2[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) + t2[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           t2[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  return o;
}

__device__ __forceinline__ BFd rot_to_xy(const BFd& b, const double* nn, const double* t1, const double* t2) {
  BFd o;
  o.v1 = nn[0] * b.v1 + t1[0] * b.v2 + t2[0] * b.v3;
  o.v2 = nn[1] * b.v1 + t1[1] * b.v2 + t2[1] * b.v3;
  o.v3 = nn[2] * b.v1 + t1[2] * b.v2 + t2[2] * b.v3;
  o.s11 = (nn[0] * (b.s11 * nn[0] + b.s12 * t1[0] + b.s13 * t2[0]) + t1[0] * (b.s12 * nn[0] + b.s22 * t1[0] + b.s23 * t2[0]) +
           t2[0] * (b.s13 * nn[0] + b.s23 * t1[0] + b.s33 * t2[0]));
  o.s12 = (nn[0] * (b.s11 * nn[1] + b.s12 * t1[1] + b.s13 * t2[1]) + t1[0] * (b.s12 * nn[1] + b.s22 * t1[1] + b.s23 * t2[1]) +
           t2[0] * (b.s13 * nn[1] + b.s23 * t1[1] + b.s33 * t2[1]));
  o.s13 = (nn[0] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) + t1[0] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[0] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  o.s22 = (nn[1] * (b.s11 * nn[1] + b.s12 * t1[1] + b.s13 * t2[1]) + t1[1] * (b.s12 * nn[1] + b.s22 * t1[1] + b.s23 * t2[1]) +
           t2[1] * (b.s13 * nn[1] + b.s23 * t1[1] + b.s33 * t2[1]));
  o.s23 = (nn[1] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) + t1[1] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[1] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  o.s33 = (nn[2] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) + t1[2] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[2] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  return o;
}

template <int M, typename R>
__device__ __forceinline__ BFd gather_bf(const R* __restrict__ f, long long N, long long id) {
  BFd b;
  if constexpr (M == 2) {
    b.v1 = (double)f[id]; b.v2 = (double)f[N + id]; b.v3 = (double)f[2 * N + id];
    b.s11 = (double)f[3 * N + id]; b.s12 = (double)f[4 * N + id]; b.s13 = (double)f[5 * N + id];
    b.s22 = (double)f[6 * N + id]; b.s23 = (double)f[7 * N + id]; b.s33 = (double)f[8 * N + id];
  } else if constexpr (M == 0) {
    b.v1 = (double)f[id]; b.v2 = (double)f[N + id]; b.v3 = 0.;
    b.s11 = (double)f[2 * N + id]; b.s12 = (double)f[3 * N + id]; b.s13 = 0.;
    b.s22 = (double)f[4 * N + id]; b.s23 = 0.; b.s33 = 0.;
  } else {
    b.v1 = 0.; b.v2 = 0.; b.v3 = (double)f[id];
    b.s11 = 0.; b.s12 = 0.; b.s13 = (double)f[N + id]; b.s22 = 0.; b.s23 = (double)f[2 * N + id]; b.s33 = 0.;
  }
  return b;
}

// ---------------------------------------------------------------------------
// bulk stage kernel

template <int M>
struct Nf {
  static constexpr int v = (M == 0) ? 5 : ((M == 1) ? 3 : 9);
};

// One stencil tap in direction d: neighbour i2 with coefficient c.  n[] is
// the metric row d at the point itself (stress terms use it), the velocity
// terms use the metric row and Jacobian at the neighbour.  Expressions
// restate block.cpp:1067-1081 (mode II), 1402-1405 (mode III), 1563-1594 (3D).
template <int M, typename R, typename C>
__device__ __forceinline__ void tap_terms(C* acc, const R* __restrict__ f, const R* __restrict__ mt,
                                          const R* __restrict__ jac, long long N, long long i2, C c, bool minus,
                                          int d, const C* n, C g2lam, C g, C lam, C invjac) {
  const C J2 = (C)jac[i2];
  const C w = invjac * c * J2;
  if constexpr (M == 0) {
    const C m0 = (C)mt[(2 * d + 0) * N + i2], m1 = (C)mt[(2 * d + 1) * N + i2];
    const C vx = (C)f[i2], vy = (C)f[N + i2], sxx = (C)f[2 * N + i2], sxy = (C)f[3 * N + i2], syy = (C)f[4 * N + i2];
    C t[5];
    t[0] = w * (m0 * sxx + m1 * sxy);
    t[1] = w * (m0 * sxy + m1 * syy);
    t[2] = c * (g2lam * n[0] * vx + lam * n[1] * vy);
    t[3] = g * c * (n[0] * vy + n[1] * vx);
    t[4] = c * (g2lam * n[1] * vy + lam * n[0] * vx);
    if (minus) {
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] -= t[q];
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] += t[q];
    }
  } else if constexpr (M == 1) {
    const C m0 = (C)mt[(2 * d + 0) * N + i2], m1 = (C)mt[(2 * d + 1) * N + i2];
    const C vz = (C)f[i2], sxz = (C)f[N + i2], syz = (C)f[2 * N + i2];
    C t[3];
    t[0] = w * (m0 * sxz + m1 * syz);
    t[1] = g * n[0] * c * vz;
    t[2] = g * n[1] * c * vz;
    if (minus) {
#pragma unroll
      for (int q = 0; q < 3; q++) acc[q] -= t[q];
    } else {
#pragma unroll
      for (int q = 0; q < 3; q++) acc[q] += t[q];
    }
  } else {
    const C m0 = (C)mt[(3 * d + 0) * N + i2], m1 = (C)mt[(3 * d + 1) * N + i2], m2 = (C)mt[(3 * d + 2) * N + i2];
    const C vx = (C)f[i2], vy = (C)f[N + i2], vz = (C)f[2 * N + i2];
    const C sxx = (C)f[3 * N + i2], sxy = (C)f[4 * N + i2], sxz = (C)f[5 * N + i2];
    const C syy = (C)f[6 * N + i2], syz = (C)f[7 * N + i2], szz = (C)f[8 * N + i2];
    C t[9];
    t[0] = w * (m0 * sxx + m1 * sxy + m2 * sxz);
    t[1] = w * (m0 * sxy + m1 * syy + m2 * syz);
    t[2] = w * (m0 * sxz + m1 * syz + m2 * szz);
    t[3] = c * (g2lam * n[0] * vx + lam * (n[1] * vy + n[2] * vz));
    t[4] = c * g * (n[0] * vy + n[1] * vx);
    t[5] = c * g * (n[0] * vz + n[2] * vx);
    t[6] = c * (g2lam * n[1] * vy + lam * (n[0] * vx + n[2] * vz));
    t[7] = c * g * (n[1] * vz + n[2] * vy);
    t[8] = c * (g2lam * n[2] * vz + lam * (n[0] * vx + n[1] * vy));
    if (minus) {
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q] -= t[q];
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q] += t[q];
    }
  }
}

// external boundary SAT of one face at one point, boundary.cpp:265-492.
// Fault-frame algebra in FP64 whatever the bulk precision.
template <int M, typename R, typename C>
__device__ void ext_sat(C* acc, const FaceDev& F, const R* __restrict__ f, const R* __restrict__ mat,
                        const GridDev& g, long long id, int u, int v, double dt) {
  const long long N = g.N;
  const long long nface = (long long)F.n0 * F.n1;
  const long long o = (long long)u * F.n1 + v;
  const int ndim = (M == 2) ? 3 : 2;
  double nn[3] = {0., 0., 0.}, t1[3], t2[3];
  for (int k = 0; k < ndim; k++) nn[k] = F.nrm[k * nface + o];
  const double h = dt * F.dl[o];
  double cp = F.cp, cs = F.cs, zp = F.zp, zs = F.zs, gamma = F.gamma;
  if (g.hetmat) {
    const double rho = (double)mat[id];
    if constexpr (M == 1) {
      cs = sqrt((double)mat[N + id] / rho);
      zs = rho * cs;
    } else {
      const double lam = (double)mat[N + id], G = (double)mat[2 * N + id];
      cp = sqrt((lam + 2. * G) / rho);
      cs = sqrt(G / rho);
      zp = rho * cp;
      zs = rho * cs;
      const double q = cs / cp;
      gamma = 1. - 2. * (q * q);
    }
  }
  if (fabs(nn[0]) > fabs(nn[1]) && fabs(nn[0]) > fabs(nn[2])) {  // boundary.cpp:316-331
    t1[2] = 0.;
    t1[1] = nn[0] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
    t1[0] = -nn[1] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
  } else if (fabs(nn[1]) > fabs(nn[2])) {
    t1[2] = 0.;
    t1[0] = nn[1] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
    t1[1] = -nn[0] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
  } else {
    t1[1] = 0.;
    t1[0] = nn[2] / sqrt(nn[0] * nn[0] + nn[2] * nn[2]);
    t1[2] = -nn[0] / sqrt(nn[0] * nn[0] + nn[2] * nn[2]);
  }
  t2[0] = nn[1] * t1[2] - nn[2] * t1[1];
  t2[1] = nn[2] * t1[0] - nn[0] * t1[2];
  t2[2] = nn[0] * t1[1] - nn[1] * t1[0];
  BFd br = rot_to_nt(gather_bf<M, R>(f, N, id), nn, t1, t2);
  BFd brs = br;
  // characteristic targets (calc_hat, boundary.cpp:494-503)
  const double ps = (F.r + 1.) * 0.5 * (br.s11 - zp * br.v1), pv = (F.r - 1.) * 0.5 * (br.s11 / zp - br.v1);
  const double s1s = (F.r + 1.) * 0.5 * (br.s12 - zs * br.v2), s1v = (F.r - 1.) * 0.5 * (br.s12 / zs - br.v2);
  const double s2s = (F.r + 1.) * 0.5 * (br.s13 - zs * br.v3), s2v = (F.r - 1.) * 0.5 * (br.s13 / zs - br.v3);
  br.v1 -= pv;
  br.v2 = 0.;
  br.v3 = 0.;
  br.s22 = gamma * (br.s11 - ps);
  br.s33 = gamma * (br.s11 - ps);
  br.s11 -= ps;
  br.s12 = 0.;
  br.s13 = 0.;
  br.s23 = 0.;
  BFd o1 = rot_to_xy(br, nn, t1, t2);
  if constexpr (M == 2) {
    const double* a = &o1.v1;
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = (C)((double)acc[q] - F.alpha * cp * h * a[q]);
#pragma unroll
    for (int q = 3; q < 9; q++) acc[q] = (C)((double)acc[q] - F.theta * cp * h * a[q]);
  } else if constexpr (M == 0) {
    acc[0] = (C)((double)acc[0] - cp * h * o1.v1);
    acc[1] = (C)((double)acc[1] - cp * h * o1.v2);
    acc[2] = (C)((double)acc[2] - cp * h * o1.s11);
    acc[3] = (C)((double)acc[3] - cp * h * o1.s12);
    acc[4] = (C)((double)acc[4] - cp * h * o1.s22);
    if (g.plastic) acc[5] = (C)((double)acc[5] - cp * h * o1.s33);
  }
  brs.v1 = 0.;
  brs.v2 -= s1v;
  brs.v3 -= s2v;
  brs.s11 = 0.;
  brs.s12 -= s1s;
  brs.s13 -= s2s;
  brs.s22 = 0.;
  brs.s23 = 0.;
  brs.s33 = 0.;
  BFd o2 = rot_to_xy(brs, nn, t1, t2);
  if constexpr (M == 2) {
    const double* a = &o2.v1;
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = (C)((double)acc[q] - F.alpha * cs * h * a[q]);
#pragma unroll
    for (int q = 3; q < 9; q++) acc[q] = (C)((double)acc[q] - F.theta * cs * h * a[q]);
  } else if constexpr (M == 0) {
    acc[0] = (C)((double)acc[0] - cs * h * o2.v1);
    acc[1] = (C)((double)acc[1] - cs * h * o2.v2);
    acc[2] = (C)((double)acc[2] - cs * h * o2.s11);
    acc[3] = (C)((double)acc[3] - cs * h * o2.s12);
    acc[4] = (C)((double)acc[4] - cs * h * o2.s22);
  } else {
    acc[0] = (C)((double)acc[0] - cs * h * o2.v3);
    acc[1] = (C)((double)acc[1] - cs * h * o2.s13);
    acc[2] = (C)((double)acc[2] - cs * h * o2.s23);
  }
}

// Fused stage update of one block: block::calc_df + set_boundaries
// (block.cpp:405-440), the interface SAT additions (interface.cpp:826-962,
// precomputed by fault_kernel), and fields::update (fields.cpp:302-309).
// Per point the accumulation order is the reference's: all x taps, then y,
// then z; each tap adds every field term then the dissipation terms.
template <int M, typename R, typename C, bool DISS, bool HET>
__global__ void __launch_bounds__(256) bulk_kernel(BlockDev bd, Boxes bx, GridDev g, StageDev st, const R* __restrict__ f,
                                                   R* __restrict__ fn, R* __restrict__ df, const R* __restrict__ mt,
                                                   const R* __restrict__ jac, const R* __restrict__ mat,
                                                   const double* __restrict__ hs) {
  constexpr int NFE = Nf<M>::v;
  constexpr int ND = (M == 2) ? 3 : 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bx.start[bx.nbox]) return;
  int bb = 0;
  while (t >= bx.start[bb + 1]) bb++;
  const long long loc = t - bx.start[bb];
  const int lk = (int)(loc % bx.n[bb][2]);
  const long long r = loc / bx.n[bb][2];
  const int lj = (int)(r % bx.n[bb][1]);
  const int li = (int)(r / bx.n[bb][1]);
  const int c3[3] = {bx.lo[bb][0] + li, bx.lo[bb][1] + lj, bx.lo[bb][2] + lk};
  const long long N = g.N;
  const long long stride[3] = {(long long)g.Ny * g.Nz, (long long)g.Nz, 1};
  const long long id = c3[0] * stride[0] + c3[1] * stride[1] + c3[2];
  const int s = g.s;
  const int NF = g.nf;  // may include szz (plastic mode II)

  double K[3][4];
  if (HET) {
    const double rho = (double)mat[id];
    for (int d = 0; d < ND; d++) {
      if constexpr (M == 1) {
        const double G = (double)mat[N + id];
        K[d][0] = st.dt / rho / bd.dx[d];
        K[d][2] = st.dt * G / bd.dx[d];
        K[d][1] = 0.;
        K[d][3] = 0.;
      } else {
        const double lam = (double)mat[N + id], G = (double)mat[2 * N + id];
        K[d][0] = st.dt / rho / bd.dx[d];
        K[d][1] = st.dt * (2. * G + lam) / bd.dx[d];
        K[d][2] = st.dt * G / bd.dx[d];
        K[d][3] = st.dt * lam / bd.dx[d];
      }
    }
  } else {
    for (int d = 0; d < ND; d++)
      for (int q = 0; q < 4; q++) K[d][q] = bd.K[d][q];
  }

  C acc[NFE + 1];
  if (st.dfmode & DF_LOAD) {
#pragma unroll
    for (int q = 0; q < NFE; q++) acc[q] = (C)df[q * N + id] * (C)st.A;
    acc[NFE] = (NF > NFE) ? (C)df[NFE * N + id] * (C)st.A : (C)0;
  } else {
#pragma unroll
    for (int q = 0; q <= NFE; q++) acc[q] = (C)0;
  }
  const C J1 = (C)jac[id];

#pragma unroll
  for (int d = 0; d < ND; d++) {
    const C invjac = (C)K[d][0] / J1;
    C nrow[3];
#pragma unroll
    for (int m = 0; m < ND; m++) nrow[m] = (C)mt[(ND * d + m) * N + id];
    const int pos = c3[d];
    int row, ntap, start, step;
    bool minus = false;
    if (pos < bd.mc[d]) {
      row = pos - bd.lo[d] + 1;
      ntap = 3 * (s - 1);
      start = bd.lo[d];
      step = 1;
    } else if (pos < bd.mrb[d]) {
      row = 0;
      ntap = 2 * s - 1;
      start = pos - s + 1;
      step = 1;
    } else {
      row = bd.hi[d] - pos;
      ntap = 3 * (s - 1);
      start = bd.hi[d] - 1;
      step = -1;
      minus = true;
    }
    const long long base = id - (long long)pos * stride[d];
    for (int nt = 0; nt < ntap; nt++) {
      const long long i2 = base + (long long)(start + step * nt) * stride[d];
      tap_terms<M, R, C>(acc, f, mt, jac, N, i2, (C)c_fd[row][nt], minus, d, nrow, (C)K[d][1], (C)K[d][2], (C)K[d][3],
                         invjac);
      if (DISS) {
        const C dc = (C)c_diss[row][nt];
        const C J2 = (C)jac[i2];
#pragma unroll
        for (int q = 0; q < NFE; q++) acc[q] += (C)g.cdiss * (dc * J2 * (C)f[q * N + i2]) / J1;
      }
    }
  }
  if constexpr (M == 0) if (NF > NFE) {  // block::calc_df_szz, block.cpp:1280-1295 (assignment, not increment)
    double nu = bd.nu;
    if (HET) nu = 0.5 * (double)mat[N + id] / ((double)mat[N + id] + (double)mat[stride[0] + id]);
    acc[NFE] = (C)nu * (acc[2] + acc[4]);
  }
  // external boundaries of this block, in face order (block.cpp:429-440)
#pragma unroll
  for (int fc = 0; fc < 2 * ND; fc++) {
    const FaceDev& F = bd.face[fc];
    if (!F.active) continue;
    const int d = fc / 2;
    const int edge = (fc & 1) ? bd.hi[d] - 1 : bd.lo[d];
    if (c3[d] != edge) continue;
    const int a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
    ext_sat<M, R, C>(acc, F, f, mat, g, id, c3[a] - bd.lo[a], c3[b] - bd.lo[b], st.dt);
  }
  // interface SAT terms in interface order (domain.cpp:251-256)
  for (int q = 0; q < bd.nif; q++) {
    const IfFaceDev& I = bd.ifc[q];
    if (c3[I.dir] != I.fixed) continue;
    const long long o = (long long)(c3[I.a] - I.org_a) * I.n1 + (c3[I.b] - I.org_b);
    const double* sn = I.sat + (long long)(I.side * 2 + 0) * NF * I.n;
    const double* sv = I.sat + (long long)(I.side * 2 + 1) * NF * I.n;
    for (int p = 0; p < NF; p++) {
      C v = (p < NFE) ? acc[p] : acc[NFE];
      v = (C)((double)v - sn[(long long)p * I.n + o]);
      v = (C)((double)v - sv[(long long)p * I.n + o]);
      if (p < NFE) acc[p] = v; else acc[NFE] = v;
    }
  }
  if (st.dfmode & DF_STORE) {
#pragma unroll
    for (int q = 0; q < NFE; q++) df[q * N + id] = (R)acc[q];
    if (NF > NFE) df[NFE * N + id] = (R)acc[NFE];
  }
  if (!(st.dfmode & F_STORE)) return;
  // RK update into the other f buffer, plus the per-step stress round trip
  C upd[NFE + 1];
#pragma unroll
  for (int q = 0; q < NFE; q++) upd[q] = (C)f[q * N + id] + (C)st.B * acc[q];
  if (NF > NFE) upd[NFE] = (C)f[NFE * N + id] + (C)st.B * acc[NFE];
  if (st.roundtrip) {
    for (int i = 0; i < g.ns; i++) {
      const int q = g.nv + i;
      const double s0 = g.s0[g.sidx[i]];
      double v = (double)((q < NFE) ? upd[q] : upd[NFE]);
      v += s0;
      if (g.hetstress) v += hs[i * N + id];
      v -= s0;
      if (g.hetstress) v -= hs[i * N + id];
      if (q < NFE) upd[q] = (C)v; else upd[NFE] = (C)v;
    }
  }
#pragma unroll
  for (int q = 0; q < NFE; q++) fn[q * N + id] = (R)upd[q];
  if (NF > NFE) fn[NFE * N + id] = (R)upd[NFE];
}

// ---------------------------------------------------------------------------
// fault kernel

struct FaultDev {
  int law, dir, a, b;
  int n0, n1, n;
  int org[3];
  double cp1, cs1, zp1, zs1, cp2, cs2, zp2, zs2, gamma1, gamma2;
  const double* nrm;
  const double* dl1;
  const double* dl2;
  int nloads, load_file, nperts, param_file;
  const double* load_xy;   // [nloads][n]
  const double* load_par;  // [nloads][10]: t0, amplitudes
  const double* lf;        // [3][n]
  const double* pert_xy;   // [nperts][n]
  const double* pert_par;  // [nperts][10]
  const double* pf;        // [9][n]
  double *u, *du, *v, *s, *sn, *ux, *dux, *vx, *sx, *state, *dstate;
  double* sat;             // [2][2][nf][n]
  double* front;           // [n] or null
  double front_value;
  int front_field;
  // STZ root: learned proposal (MLP parameters, null = plain solve only),
  // solve log rows (phi, sn, chi, eta, log V) and counters
  // [solves, Newton iterations, proposals used, fallbacks to the plain solve]
  const float* mlp;
  double* log;
  unsigned long long* log_n;
  long long log_cap;
  unsigned long long* stats;
  int stz_exact_newton;  // 1: plain solves use the exact derivative (comparison runs only)
};

struct FaultStageDev {
  double A, B, dt, t;
  int no_sat;
  double front_t;  // < 0: no front update
  // when step != null the times come from the device step counter (graph
  // replay): t = step*dt + C*dt (domain.cpp:254), front time (step+1)*dt
  const int* step;
  double C;
  int front;
};

__device__ __forceinline__ void stage_times(const FaultStageDev& fs, double& t, double& tfront) {
  if (fs.step) {
    const double ti = (double)(*fs.step) * fs.dt;
    t = ti + fs.C * fs.dt;
    tfront = fs.front ? (double)(*fs.step + 1) * fs.dt : -1.;
  } else {
    t = fs.t;
    tfront = fs.front_t;
  }
}

__global__ void advance_step(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// value of perturbation k, quantity q, at shape value xyv = xy[k][p]
__device__ __forceinline__ double pert_valx(double xyv, const double* par, int k, int q, int n, int p, double t) {
  const double t0 = par[k * 10];
  const double tf = (t >= t0) ? 1. : t / t0;
  return par[k * 10 + 1 + q] * xyv * tf;
}

// STZ slip rate and its derivative, stz.cpp:436-484
__device__ __forceinline__ double stz_vpl(double mu, const double* p) {
  if (mu <= p[7]) return 0.;
  return p[4] * exp(-p[5] + mu / p[6] - 1. / p[3]) * (1. - p[7] / mu);
}
__device__ __forceinline__ double stz_f(double mu, const double* p) { return p[1] * stz_vpl(mu, p) - p[2] * mu - p[0]; }
// df/dmu as the reference computes it (stz.cpp:471-484): the derivative of
// the factor (1 - muy/mu) enters with a minus sign where the product rule
// gives a plus, so the plain Newton iteration converges linearly (not
// quadratically) to the same root.  Kept for the plain solve, whose iterates
// then match the reference's.
__device__ __forceinline__ double stz_df(double mu, const double* p) {
  double d = (mu <= p[7]) ? 0. : stz_vpl(mu, p) * (1. / p[6] - p[7] / mu / (mu - p[7]));
  return p[1] * d - p[2];
}
// the exact derivative, dV/dmu = V (1/a + muy / (mu (mu - muy)))
__device__ __forceinline__ double stz_df_exact(double mu, const double* p) {
  double d = (mu <= p[7]) ? 0. : stz_vpl(mu, p) * (1. / p[6] + p[7] / mu / (mu - p[7]));
  return p[1] * d - p[2];
}

// the safeguarded Newton iteration of utilities.cpp:89-180 from x with
// bracket [xl, xh] (f(xl) < 0 < f(xh)): a Newton step unless it leaves the
// bracket or converges too slowly, else bisection; stops when |dx| < 2e-16
template <bool EXACT>
__device__ __forceinline__ double stz_rtsafe(double x, double xl, double xh, const double* p, int* err, int* iters) {
  int i = 0;
  const int nmax = 100;
  const double tol = 2.e-16;
  double func = stz_f(x, p);
  double der = EXACT ? stz_df_exact(x, p) : stz_df(x, p);
  double dx = fabs(xh - xl), dxold = dx;
  while (i < nmax) {
    if (der == 0.) { *err |= 1; break; }
    if ((((x - xh) * der - func) * ((x - xl) * der - func) > 0.) || fabs(2. * func) > fabs(dxold * der)) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      x = xl + dx;
      if (xl == x) break;
    } else {
      dxold = dx;
      dx = func / der;
      const double tmp = x;
      x = x - dx;
      if (tmp == x) break;
    }
    if (fabs(dx) < tol) break;
    func = stz_f(x, p);
    der = EXACT ? stz_df_exact(x, p) : stz_df(x, p);
    if (func < 0.) xl = x; else xh = x;
    i++;
  }
  if (i == nmax) *err |= 2;
  *iters = i;
  return x;
}

// bracketed Newton with bisection fallback, utilities.cpp:89-180; err codes:
// 1 zero derivative, 2 no convergence, 4 overflow repair did not terminate.
// EXACT = false: the reference's derivative (the plain solve).
template <bool EXACT>
__device__ double stz_newton(double xmin, double xmax, const double* p, int* err, int* iters) {
  double x = 0.5 * (xmin + xmax), xh = xmax, xl = xmin;
  double func = stz_f(x, p);
  const double fh = stz_f(xh, p);
  if (isinf(fh)) {
    int guard = 0;
    while (true) {
      if (!isinf(func)) {
        if (func > 0.) { xh = x; break; }
        xl = x;
      } else {
        xh = x;
      }
      x = 0.5 * (xh + xl);
      func = stz_f(x, p);
      if (++guard > 4000) { *err |= 4; break; }
    }
  }
  x = 0.5 * (xh + xl);
  return stz_rtsafe<EXACT>(x, xl, xh, p, err, iters);
}

// The same root from a proposal x0 inside the bracket: the bracket shrinks to
// the side of x0 that keeps the sign change, then the same safeguarded
// iteration runs with the exact derivative (quadratic convergence).
// Accepted only if the result stays finite, inside the bracket, and has a
// residual |f| <= 1e-12 |phi|; otherwise *err is set and the caller runs the
// plain solve.
__device__ double stz_newton_from(double xmin, double xmax, double x0, const double* p, int* err, int* iters) {
  *iters = 0;
  if (!(x0 > xmin && x0 < xmax)) { *err |= 32; return x0; }
  const double f0 = stz_f(x0, p);
  if (!isfinite(f0)) { *err |= 32; return x0; }
  double xl = xmin, xh = xmax;
  if (f0 < 0.) xl = x0; else xh = x0;
  int e = 0;
  const double x = stz_rtsafe<true>(x0, xl, xh, p, &e, iters);
  if (e || !isfinite(x) || x < xmin || x > xmax || !(fabs(stz_f(x, p)) <= 1e-12 * fabs(p[0]))) *err |= 32;
  return x;
}

// interface::apply_bcs for one fault point (interface.cpp:573-970) with the
// friction laws of friction.cpp:259-435, slipweak.cpp:206-308 and
// stz.cpp:263-356, then friction::calc_df/update (friction.cpp:465-505).
// LAW >= 0 fixes the friction law at compile time (only that law's code is
// generated); LAW = -1 reads it from F.law.  With a Hook (the STZ proposal of
// a whole wave), every lane of the wave calls this function, lanes past the
// end with live = false and the last point: once the locked shear traction,
// normal stress and impedance are known (before any write), all lanes call
// hook(phi, sn, eta, chi) for the proposed log V, and the dead lanes return.
// Without a Hook, the network in mlp (if any) is evaluated per point.
struct NoHook {
  __device__ double operator()(double, double, double, double) const { return NAN; }
};

// LEAN: the same arithmetic in fewer registers (about 115 VGPRs against 170
// for slip weakening in 2-D), for a kernel that runs it beside the bulk
// tiles: the perturbation shape values and the fault state the update needs
// are loaded where they are used instead of up front, and the interface
// frame is formed again before the SAT terms instead of being held.
template <int M, typename R, int LAW = -1, typename Hook = NoHook, bool LEAN = false>
__device__ __forceinline__ void fault_point(const FaultDev& F, const GridDev& g, const FaultStageDev& fs,
                                            const R* __restrict__ f, const R* __restrict__ mat,
                                            const double* __restrict__ hs, int* __restrict__ err,
                                            int* __restrict__ newton_iters, int p, const float* mlp,
                                            const Hook& hook = Hook(), bool live = true) {
  const int law = LAW >= 0 ? LAW : F.law;
  // this point's fault state and perturbation shape values, loaded before any
  // other memory access so that they are in flight together (and before any
  // write of this function, which may not alias them)
  double u_p = 0., du_p = 0., ux_p[2] = {0., 0.}, dux_p[2] = {0., 0.}, state_p = 0., dstate_p = 0., front_p = 0.;
  double lxy[4] = {0., 0., 0., 0.}, pxy[4] = {0., 0., 0., 0.};
  auto load_state = [&]() {  // what the slip/state update reads
    du_p = F.du[p];
#pragma unroll
    for (int q = 0; q < ((M == 2) ? 2 : 1); q++) {
      ux_p[q] = F.ux[(long long)q * F.n + p];
      dux_p[q] = F.dux[(long long)q * F.n + p];
    }
  };
  if (law != 0) {
    const int nn_ = F.n;
    u_p = F.u[p];
    if constexpr (!LEAN) load_state();
    if (law == 3) {
      state_p = F.state[p];
      dstate_p = F.dstate[p];
    }
    if constexpr (!LEAN) {
      if (F.front && F.front_field == 0) front_p = F.front[p];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (k < F.nloads) lxy[k] = F.load_xy[(long long)k * nn_ + p];
        if (k < F.nperts) pxy[k] = F.pert_xy[(long long)k * nn_ + p];
      }
    }
  }
  // shape value of load / perturbation k at this point (preloaded for k < 4)
  auto load_x = [&](int k) {
    if constexpr (LEAN) return F.load_xy[(long long)k * F.n + p];
    else return k == 0 ? lxy[0] : k == 1 ? lxy[1] : k == 2 ? lxy[2] : k == 3 ? lxy[3] : F.load_xy[(long long)k * F.n + p];
  };
  auto pert_x = [&](int k) {
    if constexpr (LEAN) return F.pert_xy[(long long)k * F.n + p];
    else return k == 0 ? pxy[0] : k == 1 ? pxy[1] : k == 2 ? pxy[2] : k == 3 ? pxy[3] : F.pert_xy[(long long)k * F.n + p];
  };
  const int ii = p / F.n1, jj = p % F.n1;
  const long long N = g.N;
  const long long stride[3] = {(long long)g.Ny * g.Nz, (long long)g.Nz, 1};
  int c[3];
  c[F.dir] = F.org[F.dir];
  c[F.a] = F.org[F.a] + ii;
  c[F.b] = F.org[F.b] + jj;
  const long long id1 = c[0] * stride[0] + c[1] * stride[1] + c[2];
  const long long id2 = id1 + stride[F.dir];
  const double h1 = fs.dt * F.dl1[p], h2 = fs.dt * F.dl2[p];
  double cp1 = F.cp1, cs1 = F.cs1, zp1 = F.zp1, zs1 = F.zs1, cp2 = F.cp2, cs2 = F.cs2, zp2 = F.zp2, zs2 = F.zs2;
  double gamma1 = F.gamma1, gamma2 = F.gamma2;
  if (g.hetmat) {
    const double r1 = (double)mat[id1], r2 = (double)mat[id2];
    if constexpr (M == 1) {
      cs1 = sqrt((double)mat[N + id1] / r1); zs1 = r1 * cs1;
      cs2 = sqrt((double)mat[N + id2] / r2); zs2 = r2 * cs2;
    } else {
      cp1 = sqrt(((double)mat[N + id1] + 2. * (double)mat[2 * N + id1]) / r1);
      cs1 = sqrt((double)mat[2 * N + id1] / r1);
      zp1 = r1 * cp1; zs1 = r1 * cs1;
      double q = cs1 / cp1; gamma1 = 1. - 2. * (q * q);
      cp2 = sqrt(((double)mat[N + id2] + 2. * (double)mat[2 * N + id2]) / r2);
      cs2 = sqrt((double)mat[2 * N + id2] / r2);
      zp2 = r2 * cp2; zs2 = r2 * cs2;
      q = cs2 / cp2; gamma2 = 1. - 2. * (q * q);
    }
  }
  // the interface frame at this point: unit normal nn, tangents t1 and t2
  // (interface.cpp:647-686)
  auto frame = [&](double* nn, double* t1, double* t2) {
    nn[0] = nn[1] = nn[2] = 0.;
    const int ndim = (M == 2) ? 3 : 2;
    for (int k = 0; k < ndim; k++) nn[k] = F.nrm[(long long)k * F.n + p];
    if (F.dir == 0) {
      t1[2] = 0.;
      t1[1] = nn[0] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
      t1[0] = -nn[1] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
    } else if (F.dir == 1) {
      t1[2] = 0.;
      t1[0] = nn[1] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
      t1[1] = -nn[0] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
    } else {
      t1[1] = 0.;
      t1[0] = nn[2] / sqrt(nn[0] * nn[0] + nn[2] * nn[2]);
      t1[2] = -nn[0] / sqrt(nn[0] * nn[0] + nn[2] * nn[2]);
    }
    t2[0] = nn[1] * t1[2] - nn[2] * t1[1];
    t2[1] = nn[2] * t1[0] - nn[0] * t1[2];
    t2[2] = nn[0] * t1[1] - nn[1] * t1[0];
    if (F.dir == 1) { t2[0] = -t2[0]; t2[1] = -t2[1]; t2[2] = -t2[2]; }
  };
  double nn[3], t1[3], t2[3];
  frame(nn, t1, t2);
  BFd b1 = gather_bf<M, R>(f, N, id1), b2 = gather_bf<M, R>(f, N, id2);
  const double* s0 = g.s0;
  if constexpr (M == 2) {
    b1.s11 += s0[0]; b1.s12 += s0[1]; b1.s13 += s0[2]; b1.s22 += s0[3]; b1.s23 += s0[4]; b1.s33 += s0[5];
    b2.s11 += s0[0]; b2.s12 += s0[1]; b2.s13 += s0[2]; b2.s22 += s0[3]; b2.s23 += s0[4]; b2.s33 += s0[5];
    if (g.hetstress) {
      b1.s11 += hs[id1]; b1.s12 += hs[N + id1]; b1.s13 += hs[2 * N + id1];
      b1.s22 += hs[3 * N + id1]; b1.s23 += hs[4 * N + id1]; b1.s33 += hs[5 * N + id1];
      b2.s11 += hs[id2]; b2.s12 += hs[N + id2]; b2.s13 += hs[2 * N + id2];
      b2.s22 += hs[3 * N + id2]; b2.s23 += hs[4 * N + id2]; b2.s33 += hs[5 * N + id2];
    }
  } else if constexpr (M == 0) {
    b1.s11 += s0[0]; b1.s12 += s0[1]; b1.s22 += s0[3];
    b2.s11 += s0[0]; b2.s12 += s0[1]; b2.s22 += s0[3];
    if (g.hetstress) {
      b1.s11 += hs[id1]; b1.s12 += hs[N + id1]; b1.s22 += hs[2 * N + id1];
      b2.s11 += hs[id2]; b2.s12 += hs[N + id2]; b2.s22 += hs[2 * N + id2];
    }
  } else {
    b1.s11 = s0[0]; b1.s13 += s0[2]; b1.s22 = s0[3]; b1.s23 += s0[4];
    b2.s11 = s0[0]; b2.s13 += s0[2]; b2.s22 = s0[3]; b2.s23 += s0[4];
    if (g.hetstress) {
      b1.s13 += hs[id1]; b1.s23 += hs[N + id1];
      b2.s13 += hs[id2]; b2.s23 += hs[N + id2];
    }
  }
  BFd r1 = rot_to_nt(b1, nn, t1, t2), r2 = rot_to_nt(b2, nn, t1, t2);
  BFd rs1 = r1, rs2 = r2;
  // locked normal characteristic (interface.cpp:1017-1028)
  auto locked = [](double v1, double v2, double sa, double sb, double z1, double z2, double& vh, double& sh) {
    sh = (z2 * (sa - z1 * v1) + z1 * (sb + z2 * v2)) / (z1 + z2);
    vh = (sh - sa) / z1 + v1;
  };
  double hv11, hs11, hv12, hs12, hv13, hs13, hv22, hs22, hv23, hs23;
  locked(r1.v1, r2.v1, r1.s11, r2.s11, zp1, zp2, hv11, hs11);
  const double hv21 = hv11, hs21 = hs11;
  if (law == 0) {
    locked(r1.v2, r2.v2, r1.s12, r2.s12, zs1, zs2, hv12, hs12);
    locked(r1.v3, r2.v3, r1.s13, r2.s13, zs1, zs2, hv13, hs13);
    hv22 = hv12; hs22 = hs12; hv23 = hv13; hs23 = hs13;
  } else {
    double snc = hs11;
    double s12 = r1.s12, s22 = r2.s12, s13 = r1.s13, s23 = r2.s13;
    const double v12 = r1.v2, v22 = r2.v2, v13 = r1.v3, v23 = r2.v3;
    const double z1 = zs1, z2 = zs2;
    const double eta = z1 * z2 / (z1 + z2);
    double t, tfront;
    stage_times(fs, t, tfront);
    const int n = F.n;
    if (F.load_file) {
      snc += F.lf[p];
      s12 += F.lf[n + p]; s22 += F.lf[n + p];
      s13 += F.lf[2 * n + p]; s23 += F.lf[2 * n + p];
    }
    for (int k = 0; k < F.nloads; k++) {
      snc += pert_valx(load_x(k), F.load_par, k, 0, n, p, t);
      const double l2 = pert_valx(load_x(k), F.load_par, k, 1, n, p, t);
      const double l3 = pert_valx(load_x(k), F.load_par, k, 2, n, p, t);
      s12 += l2; s22 += l2; s13 += l3; s23 += l3;
    }
    const double phi2 = eta * (s12 / z1 - v12 + s22 / z2 + v22);
    const double phi3 = eta * (s13 / z1 - v13 + s23 / z2 + v23);
    const double phi = sqrt(phi2 * phi2 + phi3 * phi3);
    double lv_in = NAN;
    if constexpr (!std::is_same<Hook, NoHook>::value) {
      lv_in = hook(phi, snc, eta, state_p);
      if (!live) return;
    }
    double V, S, dstatedt = 0.;
    if (law == 2) {  // slipweak::solve_fs / calc_mu
      double c0t = 0.;
      for (int k = 0; k < F.nperts; k++) c0t += pert_valx(pert_x(k), F.pert_par, k, 3, n, p, t);
      if (F.param_file) c0t += F.pf[3 * n + p];
      if (snc < 0.) {
        double dct = 0., must = 0., mudt = 0., trupt = 0., tct = 0.;
        for (int k = 0; k < F.nperts; k++) {
          dct += pert_valx(pert_x(k), F.pert_par, k, 0, n, p, t);
          must += pert_valx(pert_x(k), F.pert_par, k, 1, n, p, t);
          mudt += pert_valx(pert_x(k), F.pert_par, k, 2, n, p, t);
          trupt += pert_valx(pert_x(k), F.pert_par, k, 4, n, p, t);
          tct += pert_valx(pert_x(k), F.pert_par, k, 5, n, p, t);
        }
        if (F.param_file) {
          dct += F.pf[p]; must += F.pf[n + p]; mudt += F.pf[2 * n + p];
          trupt += F.pf[4 * n + p]; tct += F.pf[5 * n + p];
        }
        const double u = u_p;
        const double f1 = (dct == 0. || u >= dct) ? 1. : u / dct;
        double f2;
        if (trupt <= 0. || t < trupt) f2 = 0.;
        else if (t >= trupt && t < trupt + tct) f2 = (tct == 0.) ? 1. : (t - trupt) / tct;
        else f2 = 1.;
        const double mu = (f1 >= f2) ? must + (mudt - must) * f1 : must + (mudt - must) * f2;
        if (c0t + mu * fabs(snc) > phi) { V = 0.; S = phi; }
        else { S = c0t + mu * fabs(snc); V = (phi - S) / eta; }
      } else {
        if (phi > c0t) { S = c0t; V = (phi - S) / eta; }
        else { V = 0.; S = phi; }
      }
    } else {  // friction::solve_fs with frictionless (mu = 0) or stz calc_mu
      if (snc < 0.) {
        double mu = 0.;
        if (law == 3) {
          double pr[8] = {phi, eta, snc, state_p, 0., 0., 0., 0.};
          for (int k = 0; k < F.nperts; k++) {
            pr[4] += pert_valx(pert_x(k), F.pert_par, k, 0, n, p, t);
            pr[5] += pert_valx(pert_x(k), F.pert_par, k, 1, n, p, t);
            pr[6] += pert_valx(pert_x(k), F.pert_par, k, 2, n, p, t);
            pr[7] += pert_valx(pert_x(k), F.pert_par, k, 3, n, p, t);
          }
          if (F.param_file) { pr[4] += F.pf[p]; pr[5] += F.pf[n + p]; pr[6] += F.pf[2 * n + p]; pr[7] += F.pf[3 * n + p]; }
          if (!(pr[3] > 0.)) atomicOr(err, 8);
          if (pr[7] >= -phi / snc) {
            mu = -phi / snc;
          } else {
            int e = 0, it = 0, used = 0, fell = 0;
            if (mlp) {
              // learned proposal for log V; the friction coefficient it implies,
              // from the force balance mu = (phi - eta V)/|sn| or from the
              // inverted flow law mu = a (log(V/v0) + f0 + 1/chi - log(1 - muy/mu))
              // (stz.cpp:436-484), whichever leaves the smaller residual
              const double lv = isnan(lv_in) ? mlp_logv(mlp, phi, snc, pr[3], eta) : lv_in;
              const double xmin = pr[7], xmax = -phi / snc;
              const double ma = (phi - eta * exp(lv)) / (-snc);
              const double lb = lv - log(pr[4]) + pr[5] + 1. / pr[3];
              double mb = pr[6] * lb;
              for (int q = 0; q < 2; q++)
                if (mb > xmin) mb = pr[6] * (lb - log1p(-xmin / mb));
              const bool ina = ma > xmin && ma < xmax, inb = mb > xmin && mb < xmax;
              const double ra = ina ? fabs(stz_f(ma, pr)) : INFINITY, rb = inb ? fabs(stz_f(mb, pr)) : INFINITY;
              const double mu0 = rb < ra ? mb : ma;
              int e1 = 0;
              mu = stz_newton_from(xmin, xmax, mu0, pr, &e1, &it);
              used = !e1;
              fell = e1 ? 1 : 0;
            }
            if (!used) {
              int it2 = 0;
              mu = F.stz_exact_newton ? stz_newton<true>(pr[7], -phi / snc, pr, &e, &it2)
                                      : stz_newton<false>(pr[7], -phi / snc, pr, &e, &it2);
              it += it2;
            }
            if (e) atomicOr(err, e);
            if (newton_iters) atomicAdd(newton_iters, it);
            if (F.stats) {
              atomicAdd(&F.stats[0], 1ULL);
              atomicAdd(&F.stats[1], (unsigned long long)it);
              if (used) atomicAdd(&F.stats[2], 1ULL);
              if (fell) atomicAdd(&F.stats[3], 1ULL);
            }
            if (F.log) {  // training pair of the plain method's root
              const unsigned long long r = atomicAdd(F.log_n, 1ULL);
              if ((long long)r < F.log_cap) {
                double* row = F.log + 5 * r;
                row[0] = phi;
                row[1] = snc;
                row[2] = pr[3];
                row[3] = eta;
                row[4] = log(stz_vpl(mu, pr));
              }
            }
          }
        }
        if (mu * fabs(snc) > phi) { V = 0.; S = phi; }
        else { S = mu * fabs(snc); V = (phi - S) / eta; }
      } else {
        S = 0.;
        V = phi / eta;
      }
      if (law == 3) {  // stz::calc_dstatedt / chihat
        double c0t = 0., Rt = 0., betat = 0., chiwt = 0., v1t = 0.;
        for (int k = 0; k < F.nperts; k++) {
          c0t += pert_valx(pert_x(k), F.pert_par, k, 4, n, p, t);
          Rt += pert_valx(pert_x(k), F.pert_par, k, 5, n, p, t);
          betat += pert_valx(pert_x(k), F.pert_par, k, 6, n, p, t);
          chiwt += pert_valx(pert_x(k), F.pert_par, k, 7, n, p, t);
          v1t += pert_valx(pert_x(k), F.pert_par, k, 8, n, p, t);
        }
        if (F.param_file) {
          c0t += F.pf[4 * n + p]; Rt += F.pf[5 * n + p]; betat += F.pf[6 * n + p];
          chiwt += F.pf[7 * n + p]; v1t += F.pf[8 * n + p];
        }
        if (V > v1t) atomicOr(err, 16);
        const double chihat = (V == 0.) ? chiwt : chiwt / log(v1t / V);
        dstatedt = V * S / c0t * (1. - state_p / chihat) - Rt * exp(-betat / state_p);
      }
    }
    double vv2, vv3;
    if (V == 0.) { vv2 = 0.; vv3 = 0.; }
    else { vv2 = V * phi2 / (eta * V + S); vv3 = V * phi3 / (eta * V + S); }
    hs12 = phi2 - eta * vv2;
    hs22 = hs12;
    hs13 = phi3 - eta * vv3;
    hs23 = hs13;
    hv12 = (hs12 - s12) / z1 + v12;
    hv22 = (-hs22 + s22) / z2 + v22;
    hv13 = (hs13 - s13) / z1 + v13;
    hv23 = (-hs23 + s23) / z2 + v23;
    F.v[p] = V;
    F.s[p] = S;
    F.sn[p] = snc;
    if constexpr (M == 2) {
      F.vx[p] = vv2; F.vx[n + p] = vv3;
      F.sx[p] = hs12; F.sx[n + p] = hs13;
    } else if constexpr (M == 0) {
      F.vx[p] = vv2; F.sx[p] = hs12;
    } else {
      F.vx[p] = vv3; F.sx[p] = hs13;
    }
    if (F.load_file) {
      hs12 -= F.lf[n + p]; hs22 -= F.lf[n + p];
      hs13 -= F.lf[2 * n + p]; hs23 -= F.lf[2 * n + p];
    }
    for (int k = 0; k < F.nloads; k++) {
      const double l2 = pert_valx(load_x(k), F.load_par, k, 1, n, p, t);
      const double l3 = pert_valx(load_x(k), F.load_par, k, 2, n, p, t);
      hs12 -= l2; hs22 -= l2; hs13 -= l3; hs23 -= l3;
    }
    if (!fs.no_sat) {  // friction::scale_df + calc_df + update for this point
      if constexpr (LEAN) load_state();
      double du = du_p * fs.A;
      du += fs.dt * V;
      F.du[p] = du;
      F.u[p] = u_p + fs.B * du;
      const int nc = (M == 2) ? 2 : 1;
      const double vxl[2] = {(M == 1) ? vv3 : vv2, vv3};  // the F.vx values written above
#pragma unroll
      for (int q = 0; q < nc; q++) {
        const long long o = (long long)q * n + p;
        double dux = dux_p[q] * fs.A;
        dux += fs.dt * vxl[q];
        F.dux[o] = dux;
        F.ux[o] = ux_p[q] + fs.B * dux;
      }
      if (law == 3) {
        double dst = dstate_p * fs.A;
        dst += fs.dt * dstatedt;
        F.dstate[p] = dst;
        F.state[p] = state_p + fs.B * dst;
      }
    }
    if constexpr (LEAN) {
      if (F.front && tfront >= 0. && F.front_field == 0 && V >= F.front_value && F.front[p] < 0.) F.front[p] = tfront;
    } else {
      if (F.front && tfront >= 0. && F.front_field == 0 && front_p < 0. && V >= F.front_value) F.front[p] = tfront;
    }
  }
  if (fs.no_sat) return;
  // SAT terms for both sides (interface.cpp:802-962)
  const int NF = g.nf;
  const long long n = F.n;
  double* sn1 = F.sat + (0 * 2 + 0) * NF * n;
  double* sv1 = F.sat + (0 * 2 + 1) * NF * n;
  double* sn2 = F.sat + (1 * 2 + 0) * NF * n;
  double* sv2 = F.sat + (1 * 2 + 1) * NF * n;
  r1.v1 -= hv11; r1.v2 = 0.; r1.v3 = 0.;
  r1.s22 = gamma1 * (r1.s11 - hs11);
  r1.s33 = gamma1 * (r1.s11 - hs11);
  r1.s11 -= hs11; r1.s12 = 0.; r1.s13 = 0.; r1.s23 = 0.;
  r2.v1 -= hv21; r2.v2 = 0.; r2.v3 = 0.;
  r2.s22 = gamma2 * (r2.s11 - hs21);
  r2.s33 = gamma2 * (r2.s11 - hs21);
  r2.s11 -= hs21; r2.s12 = 0.; r2.s13 = 0.; r2.s23 = 0.;
  if constexpr (LEAN) {  // formed again rather than held through the friction solve
    __builtin_amdgcn_sched_barrier(0);
    frame(nn, t1, t2);
  }
  BFd o1 = rot_to_xy(r1, nn, t1, t2), o2 = rot_to_xy(r2, nn, t1, t2);
  if constexpr (M == 2) {
    const double* a1 = &o1.v1;
    const double* a2 = &o2.v1;
    for (int q = 0; q < 9; q++) { sn1[q * n + p] = cp1 * h1 * a1[q]; sn2[q * n + p] = cp2 * h2 * a2[q]; }
  } else if constexpr (M == 0) {
    sn1[p] = cp1 * h1 * o1.v1; sn1[n + p] = cp1 * h1 * o1.v2; sn1[2 * n + p] = cp1 * h1 * o1.s11;
    sn1[3 * n + p] = cp1 * h1 * o1.s12; sn1[4 * n + p] = cp1 * h1 * o1.s22;
    sn2[p] = cp2 * h2 * o2.v1; sn2[n + p] = cp2 * h2 * o2.v2; sn2[2 * n + p] = cp2 * h2 * o2.s11;
    sn2[3 * n + p] = cp2 * h2 * o2.s12; sn2[4 * n + p] = cp2 * h2 * o2.s22;
    if (NF > 5) { sn1[5 * n + p] = cp1 * h1 * o1.s33; sn2[5 * n + p] = cp2 * h2 * o2.s33; }
  } else {
    for (int q = 0; q < 3; q++) { sn1[q * n + p] = 0.; sn2[q * n + p] = 0.; }
  }
  rs1.v1 = 0.; rs1.v2 -= hv12; rs1.v3 -= hv13;
  rs1.s11 = 0.; rs1.s12 -= hs12; rs1.s13 -= hs13; rs1.s22 = 0.; rs1.s23 = 0.; rs1.s33 = 0.;
  rs2.v1 = 0.; rs2.v2 -= hv22; rs2.v3 -= hv23;
  rs2.s11 = 0.; rs2.s12 -= hs22; rs2.s13 -= hs23; rs2.s22 = 0.; rs2.s23 = 0.; rs2.s33 = 0.;
  o1 = rot_to_xy(rs1, nn, t1, t2);
  o2 = rot_to_xy(rs2, nn, t1, t2);
  if constexpr (M == 2) {
    const double* a1 = &o1.v1;
    const double* a2 = &o2.v1;
    for (int q = 0; q < 9; q++) { sv1[q * n + p] = cs1 * h1 * a1[q]; sv2[q * n + p] = cs2 * h2 * a2[q]; }
  } else if constexpr (M == 0) {
    sv1[p] = cs1 * h1 * o1.v1; sv1[n + p] = cs1 * h1 * o1.v2; sv1[2 * n + p] = cs1 * h1 * o1.s11;
    sv1[3 * n + p] = cs1 * h1 * o1.s12; sv1[4 * n + p] = cs1 * h1 * o1.s22;
    sv2[p] = cs2 * h2 * o2.v1; sv2[n + p] = cs2 * h2 * o2.v2; sv2[2 * n + p] = cs2 * h2 * o2.s11;
    sv2[3 * n + p] = cs2 * h2 * o2.s12; sv2[4 * n + p] = cs2 * h2 * o2.s22;
    if (NF > 5) { sv1[5 * n + p] = 0.; sv2[5 * n + p] = 0.; }
  } else {
    sv1[p] = cs1 * h1 * o1.v3; sv1[n + p] = cs1 * h1 * o1.s13; sv1[2 * n + p] = cs1 * h1 * o1.s23;
    sv2[p] = cs2 * h2 * o2.v3; sv2[n + p] = cs2 * h2 * o2.s13; sv2[2 * n + p] = cs2 * h2 * o2.s23;
  }
}

// one thread per fault point; LAW is the interface's friction law (0 locked,
// 1 frictionless, 2 slip weakening, 3 STZ)
template <int M, typename R, int LAW>
__global__ void __launch_bounds__(256) fault_kernel(FaultDev F, GridDev g, FaultStageDev fs, const R* __restrict__ f,
                                                    const R* __restrict__ mat, const double* __restrict__ hs,
                                                    int* __restrict__ err, int* __restrict__ newton_iters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (LAW == 3) {
    if (F.mlp) {
      // the proposal network runs for whole waves on the matrix cores, from
      // LDS (one copy per workgroup): its loads are issued first and land in
      // LDS once the points' inputs are known, so they overlap the fault
      // point's own loads (256 threads, 4 full waves; MLP_NP <= 18 * 256)
      __shared__ float smlp[MLP_NP];
      constexpr int PER = (MLP_NP + 255) / 256;
      float wreg[PER];
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const int i = threadIdx.x + 256 * u;
        wreg[u] = i < MLP_NP ? F.mlp[i] : 0.f;
      }
      auto hook = [&](double phi, double sn, double eta, double chi) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
          const int i = threadIdx.x + 256 * u;
          if (i < MLP_NP) smlp[i] = wreg[u];
        }
        __syncthreads();
        float x[MLP_IN];
        mlp_inputs(smlp, phi, sn, chi, eta, x);
        return mlp_out(smlp, mlp_eval_wave(smlp, x));
      };
      const bool live = p < F.n;
      fault_point<M, R, 3>(F, g, fs, f, mat, hs, err, newton_iters, live ? p : F.n - 1, smlp, hook, live);
      return;
    }
    if (p >= F.n) return;
    fault_point<M, R, 3>(F, g, fs, f, mat, hs, err, newton_iters, p, nullptr);
  } else {
    if (p >= F.n) return;
    fault_point<M, R, LAW>(F, g, fs, f, mat, hs, err, newton_iters, p, nullptr);
  }
}

// ---------------------------------------------------------------------------
// plasticity: fields::set_stress, block::calc_plastic (block.cpp:442-697),
// fields::remove_stress, then the output round trip of problem.cpp:160-164.
template <int M, typename R>
__global__ void __launch_bounds__(256) plastic_kernel(BlockDev bd, GridDev g, double dt, R* __restrict__ f,
                                                      R* __restrict__ fp, const R* __restrict__ mat,
                                                      const double* __restrict__ hs) {
  const long long npts = (long long)bd.n[0] * bd.n[1] * bd.n[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const int lk = (int)(t % bd.n[2]);
  const long long r = t / bd.n[2];
  const int lj = (int)(r % bd.n[1]);
  const int li = (int)(r / bd.n[1]);
  const long long N = g.N;
  const long long id = ((long long)(bd.lo[0] + li) * g.Ny + (bd.lo[1] + lj)) * g.Nz + (bd.lo[2] + lk);
  double sv[6];
  for (int i = 0; i < g.ns; i++) {  // set_stress
    double v = (double)f[(g.nv + i) * N + id];
    v += g.s0[g.sidx[i]];
    if (g.hetstress) v += hs[i * N + id];
    sv[i] = v;
  }
  double sxx, sxy, sxz, syy, syz, szz;
  if constexpr (M == 2) { sxx = sv[0]; sxy = sv[1]; sxz = sv[2]; syy = sv[3]; syz = sv[4]; szz = sv[5]; }
  else if constexpr (M == 0) { sxx = sv[0]; sxy = sv[1]; sxz = 0.; syy = sv[2]; syz = 0.; szz = sv[3]; }
  else { sxx = g.s0[0]; sxy = 0.; sxz = sv[0]; syy = g.s0[3]; syz = sv[1]; szz = g.s0[5]; }
  double lambda = (double)fp[id], gammap = (double)fp[N + id];
  double ep[6] = {0., 0., 0., 0., 0., 0.};
  if (g.ptensor)
    for (int q = 0; q < 6; q++) ep[q] = (double)fp[(2 + q) * N + id];
  double k, G;
  if (g.hetmat) {
    k = (double)mat[N + id] + 2. / 3. * (double)mat[2 * N + id];
    G = (double)mat[2 * N + id];
  } else {
    k = bd.lam + 2. / 3. * bd.G;
    G = bd.G;
  }
  const double mu = bd.pmu, coh = bd.pc, beta = bd.pbeta, eta = bd.peta;
  double sigma = (sxx + syy + szz) / 3.;
  double tau = sqrt(((sxx - syy) * (sxx - syy) + (syy - szz) * (syy - szz) + (szz - sxx) * (szz - sxx)) / 6. + sxy * sxy +
                    sxz * sxz + syz * syz);
  double yf = coh - mu * sigma;
  if (yf < 0.) yf = 0.;
  if (tau - yf <= 0.) {
    lambda = 0.;
  } else {
    if (tau * mu * beta * k > (mu * sigma - coh) * G) lambda = (tau + mu * sigma - coh) / (eta + dt * (G + mu * beta * k));
    else lambda = tau / (eta + dt * G);
    gammap = gammap + dt * lambda;
    double sd[6] = {sxx - sigma, sxy, sxz, syy - sigma, syz, szz - sigma};
    tau -= dt * lambda * G;
    sigma -= dt * lambda * beta * k;
    for (int i = 0; i < 6; i++) sd[i] *= tau / (tau + dt * lambda * G);
    sxx = sd[0] + sigma; sxy = sd[1]; sxz = sd[2]; syy = sd[3] + sigma; syz = sd[4]; szz = sd[5] + sigma;
    ep[0] = ep[0] + dt * lambda * (sd[0] / (2. * tau) + (beta / 3.));
    ep[1] = ep[1] + dt * lambda * (sd[1] / (2. * tau));
    ep[2] = ep[2] + dt * lambda * (sd[2] / (2. * tau));
    ep[3] = ep[3] + dt * lambda * (sd[3] / (2. * tau) + (beta / 3.));
    ep[4] = ep[4] + dt * lambda * (sd[4] / (2. * tau));
    ep[5] = ep[5] + dt * lambda * (sd[5] / (2. * tau) + (beta / 3.));
  }
  if constexpr (M == 2) { sv[0] = sxx; sv[1] = sxy; sv[2] = sxz; sv[3] = syy; sv[4] = syz; sv[5] = szz; }
  else if constexpr (M == 0) { sv[0] = sxx; sv[1] = sxy; sv[2] = syy; sv[3] = szz; }
  else { sv[0] = sxz; sv[1] = syz; }
  for (int i = 0; i < g.ns; i++) {
    double v = sv[i];
    const double s0 = g.s0[g.sidx[i]];
    v -= s0;  // remove_stress
    if (g.hetstress) v -= hs[i * N + id];
    v += s0;  // set_stress / remove_stress around the output write
    if (g.hetstress) v += hs[i * N + id];
    v -= s0;
    if (g.hetstress) v -= hs[i * N + id];
    f[(g.nv + i) * N + id] = (R)v;
  }
  fp[id] = (R)lambda;
  fp[N + id] = (R)gammap;
  if (g.ptensor)
    for (int q = 0; q < 6; q++) fp[(2 + q) * N + id] = (R)ep[q];
}

// ---------------------------------------------------------------------------
// gathers for read-back and output units

// box of one field (optionally absolute stress: (f + s0) + s) into dst (FP64)
template <typename R>
__global__ void gather_box(const R* __restrict__ src, const double* __restrict__ hs, double s0add, int hs_field,
                           long long N, int Ny, int Nz, int lo0, int lo1, int lo2, int st0, int st1, int st2, int n0,
                           int n1, int n2, double* __restrict__ dst) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long tot = (long long)n0 * n1 * n2;
  if (t >= tot) return;
  const int c = (int)(t % n2);
  const long long r = t / n2;
  const int b = (int)(r % n1);
  const int a = (int)(r / n1);
  const long long id = ((long long)(lo0 + a * st0) * Ny + (lo1 + b * st1)) * Nz + (lo2 + c * st2);
  double v = (double)src[id];
  if (s0add != 0. || hs_field >= 0) {
    v += s0add;
    if (hs_field >= 0) v += hs[hs_field * N + id];
  }
  dst[t] = v;
}

// Every output unit due at a step in one launch (blockIdx.y: the unit in
// the list): the unit's index box, C order, of a field (stress fields
// absolute: f + s0 (+ s)) or of one component of a fault array
// ([ncomp][n0][n1]; a0, a1 the grid directions along the interface, xa0, xa1
// its first index in them), into the unit's staging slot
// (outputunit.cpp:771-889).
struct OutGatherDev {
  int iface;        // 0: grid field, 1: fault array
  int field;        // grid: index into f (< nf) or, from nf, the plastic fields
  int add, hsf;     // grid: absolute stress (s0add, heterogeneous stress field hsf >= 0)
  double s0add;
  int lo[3], st[3], n[3];
  const double* arr;  // fault array; n_if points per component, row length nn1
  long long n_if;
  int nn1, comp, a0, a1, xa0, xa1;
  double* dst;      // slot 0 of the staging ring; slot k at dst + k cnt
  long long cnt;
};

struct OutGatherList {
  static constexpr int MAXU = 16;
  int n;
  int unit[MAXU], slot[MAXU];
};

template <typename R>
__global__ void __launch_bounds__(256) gather_units(const OutGatherDev* __restrict__ units, OutGatherList L,
                                                    const R* __restrict__ f, const R* __restrict__ fp, int nf,
                                                    const double* __restrict__ hs, long long N, int Ny, int Nz) {
  const OutGatherDev& U = units[L.unit[blockIdx.y]];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= U.cnt) return;
  const int c = (int)(t % U.n[2]);
  const long long r = t / U.n[2];
  const int b = (int)(r % U.n[1]);
  const int a = (int)(r / U.n[1]);
  const int cc[3] = {U.lo[0] + a * U.st[0], U.lo[1] + b * U.st[1], U.lo[2] + c * U.st[2]};
  double v;
  if (U.iface) {
    v = U.arr[(long long)U.comp * U.n_if + (long long)(cc[U.a0] - U.xa0) * U.nn1 + (cc[U.a1] - U.xa1)];
  } else {
    const long long id = ((long long)cc[0] * Ny + cc[1]) * Nz + cc[2];
    v = U.field < nf ? (double)f[(long long)U.field * N + id] : (double)fp[(long long)(U.field - nf) * N + id];
    if (U.add) {  // absolute stress, the order of fields::set_stress (fields.cpp:184-220)
      v += U.s0add;
      if (U.hsf >= 0) v += hs[U.hsf * N + id];
    }
  }
  U.dst[(long long)L.slot[blockIdx.y] * U.cnt + t] = v;
}

template <typename R>
__global__ void convert_from_double(const double* __restrict__ src, R* __restrict__ dst, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (R)src[t];
}

}  // namespace fdf

namespace fdf {

// ---------------------------------------------------------------------------
// Rectilinear 3-D block kernel.
//
// On a Cartesian block the metric is diagonal and uniform (up to rounding)
// and J is constant, so the update of block.cpp:1541-2050 reduces to
// per-direction derivatives of 6 fields scaled by block constants.  Each
// thread owns one (y,z) column of a TY x 64 tile and marches along x over a
// chunk of planes:
//   * x derivatives from a register window of the 6 fields x-derivatives need
//     (vx vy vz sxx sxy sxz) over 2r+1 planes, loaded once per plane;
//   * y and z derivatives from an LDS tile of the current plane with an r-point
//     halo (8 fields: all but sxx);
//   * one-sided SBP rows near block edges read their taps directly (cached);
//   * df = A*df + ..., f_new = f + B*df, written once;
//   * external SAT as the face's 16 non-zero linear-map entries, interface
//     SAT terms from the fault kernel's buffer.
// Not bit-identical to the reference (derivative sums are
// formed before scaling, the metric is the block constant); the FP64 result
// stays within 1e-10 of the bit-faithful path over long runs.
// Workgroup barrier that orders LDS only: __syncthreads() also waits for every
// outstanding global load, which would stall the next-plane prefetch.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct Tiled3D {
  int lo[3], hi[3];              // block box
  int mc[3], mrb[3];             // interior rows [mc, mrb) per direction
  int Ny, Nz;
  long long N;
  int cx;                        // planes per chunk
  int nzt, nyt, nxc, ntiles;     // tile grid
  double cv[3], cg2[3], cl[3], cg[3];  // dt/rho/dx*m, dt(2G+lam)/dx*m, dt lam/dx*m, dt G/dx*m
  double A, B;
  float fcv[3], fcg2[3], fcl[3], fcg[3], fA, fB;  // the same, rounded once for the FP32 path
  int roundtrip, hetstress;
  int dfmode = DF_ALL;           // DF_LOAD / DF_STORE / F_STORE bits of this launch
  double s0[6];
  int rows_nht;                  // > 0: row-tile kernel with this many halo entries per thread
  int rows_var;                  // row-tile variant: 1 double-buffered LDS tile, 0 single
  int rows_shape;                // row-tile workgroup (FP32): 0 TY = 4 rows, 1 TY = 6 (LDS stride 160), 2 TY = 5
  int grp_y, grp_x;              // row-tile order: patches of grp_y y tiles x grp_x x chunks
  int wide;                      // 1: one buffer descriptor per field (9 N elements above 4 GiB)
  int face_active[6];            // external SAT faces
  // external SAT of a Cartesian face normal to axis d (other axes a < b) as
  // its 16 non-zero linear-map entries, [6 faces][16], in the bulk precision,
  // in device memory: rows (v_d, s_dd, s_aa, s_bb) x cols (v_d, s_dd), rows
  // (v_a, s_da) x cols (v_a, s_da), same for b
  const void* fcoef;
  int nif;                       // interface faces touching the block
  int if_dir[6], if_fixed[6], if_a[6], if_b[6], if_orga[6], if_orgb[6], if_n[6], if_n1[6], if_side[6];
  const double* if_sat[6];
};

// Raw buffer access (gfx9 buffer descriptor in scalar registers, 32-bit
// per-lane byte offset plus a wave-uniform scalar byte offset): no per-lane
// 64-bit address arithmetic, and no address registers for the compiler to
// hoist out of a loop.  One descriptor per field array of `count` elements.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
template <typename T>
__device__ __forceinline__ rsrc_t buf_rsrc(const T* p, long long count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), (short)0, (int)(unsigned)(count * (long long)sizeof(T)),
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bst(rsrc_t r, T v, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
  } else {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
  }
}

// A field array [nf][N] through buffer descriptors; field q at (per-lane
// byte offset voff, wave-uniform byte offset soff) within the field.
// Fields1: one descriptor over the whole array (nf N elements within 4 GiB),
// field q at scalar offset q fb.  FieldsQ: one descriptor per field, formed
// from that field's 64-bit base (each field within 4 GiB), for arrays above
// 4 GiB.  on = false gives descriptors with no records: loads return 0 and
// stores are dropped, without a memory request, and the instruction stream
// stays the same.
template <typename R>
struct Fields1 {
  rsrc_t r;
  unsigned fb;
  __device__ __forceinline__ Fields1(const R* p, long long N, int nf, bool on)
      : r(buf_rsrc(p, on ? nf * N : 0)), fb((unsigned)(N * (long long)sizeof(R))) {}
  __device__ __forceinline__ R ld(int q, unsigned voff, unsigned soff) const { return bld<R>(r, voff, q * fb + soff); }
  __device__ __forceinline__ void st(int q, R v, unsigned voff, unsigned soff) const { bst<R>(r, v, voff, q * fb + soff); }
};
template <typename R>
struct FieldsQ {
  const R* p;
  long long N, rec;
  __device__ __forceinline__ FieldsQ(const R* p_, long long N_, int, bool on) : p(p_), N(N_), rec(on ? N_ : 0) {}
  __device__ __forceinline__ R ld(int q, unsigned voff, unsigned soff) const {
    return bld<R>(buf_rsrc(p + q * N, rec), voff, soff);
  }
  __device__ __forceinline__ void st(int q, R v, unsigned voff, unsigned soff) const {
    bst<R>(buf_rsrc(p + q * N, rec), v, voff, soff);
  }
};
template <typename R, bool WIDE> struct FieldsOf { typedef Fields1<R> type; };
template <typename R> struct FieldsOf<R, true> { typedef FieldsQ<R> type; };

// one-sided SBP row whose taps lie in one x plane: taps at element offsets
// lane0 + tap*stride from the wave-uniform plane pointer fpl; left rows use
// c_fd[pos-lo+1][n] at lo+n, right rows -c_fd[hi-pos][n] at hi-1-n
template <typename R, typename C, int S, typename FA>
__device__ __forceinline__ void edge_plane(C* out, const FA& F, unsigned plb, const int* fq, int lane0, int stride,
                                           int pos, int lo, int hi, const C (*sfd)[3 * (S - 1)]) {
  constexpr int NT = 3 * (S - 1);
  const bool left = pos - lo < 2 * (S - 1);
  const int row = left ? pos - lo + 1 : hi - pos;
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int tap = left ? lo + n : hi - 1 - n;
    const C c = sfd[row - 1][n];
    const unsigned off = (unsigned)((lane0 + tap * stride) * (int)sizeof(R));
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const C v = (C)F.ld(fq[q], off, plb);
      out[q] = n == 0 ? c * v : out[q] + c * v;
    }
  }
  if (!left)
#pragma unroll
    for (int q = 0; q < 6; q++) out[q] = -out[q];
}

// one-sided SBP row along x at the wave-uniform plane i: taps are whole planes
template <typename R, typename C, int S, typename FA>
__device__ __forceinline__ void edge_x(C* out, const FA& F, unsigned psb, unsigned colb, int i, int lo, int hi) {
  constexpr int NT = 3 * (S - 1);
  const bool left = i - lo < 2 * (S - 1);
  const int row = left ? i - lo + 1 : hi - i;
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const int tap = left ? lo + n : hi - 1 - n;
    const C c = (C)c_fd[row][n];
    const unsigned plb = (unsigned)tap * psb;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const C v = (C)F.ld(q, colb, plb);
      out[q] = n == 0 ? c * v : out[q] + c * v;
    }
  }
  if (!left)
#pragma unroll
    for (int q = 0; q < 6; q++) out[q] = -out[q];
}

// interior first derivative from 2r+1 samples u[0..2r] centred on u[r]; the
// interior SBP stencil is antisymmetric (fd.cpp:52-121), cd[n] = c_fd[0][r+n]
template <typename C, int r, typename F>
__device__ __forceinline__ C central(const C* cd, F u) {
  C s = cd[1] * (u(r + 1) - u(r - 1));
#pragma unroll
  for (int n = 2; n <= r; n++) s += cd[n] * (u(r + n) - u(r - n));
  return s;
}

// WPE: minimum waves per SIMD the register allocation must allow
template <typename R, typename C, int S, int TY, int WPE, bool WIDE = false>
__global__ void __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(WPE))) bulk_tiled3d(Tiled3D ta, const R* __restrict__ f, R* __restrict__ fn,
                                                        R* __restrict__ df, const double* __restrict__ hs) {
  // This kernel serves the FP32 path and the fast FP64 path, neither of which
  // keeps the reference's association, so fused multiply-adds are allowed.
#pragma clang fp contract(fast)
  constexpr int r = S - 1;
  constexpr int W = 2 * r + 1;
  constexpr int TZ = 64;
  constexpr int LY = TY + 2 * r, LZ = TZ + 2 * r;
  constexpr int NT = 64 * TY;
  __shared__ R tile[8][LY][LZ];  // 0 vx 1 vy 2 vz 3 sxy 4 sxz 5 syy 6 syz 7 szz
  // one-sided SBP rows in LDS: read inside the loop with per-lane rows
  // (constant memory would be hoisted out of the loop into registers)
  constexpr int NRB = 2 * (S - 1), NTB = 3 * (S - 1);
  __shared__ C sfd[NRB][NTB];
  if (threadIdx.x < NRB * NTB) sfd[threadIdx.x / NTB][threadIdx.x % NTB] = (C)c_fd[1 + threadIdx.x / NTB][threadIdx.x % NTB];
  // XCD-aware tile order: consecutive hardware workgroup ids land on
  // different XCDs, so give each XCD a contiguous range of tiles
  const int nxcd = 8;
  const int per = (ta.ntiles + nxcd - 1) / nxcd;
  const int tile_id = (blockIdx.x % nxcd) * per + blockIdx.x / nxcd;
  if (tile_id >= ta.ntiles) return;
  const int zt = tile_id % ta.nzt;
  const int rem = tile_id / ta.nzt;
  const int yt = rem % ta.nyt;
  const int xc = rem / ta.nyt;
  const int tz = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int kb = ta.lo[2] + zt * TZ, jb = ta.lo[1] + yt * TY;
  const int k = kb + tz, j = jb + ty;
  const bool active = (k < ta.hi[2]) && (j < ta.hi[1]);
  const int ib = ta.lo[0] + xc * ta.cx;
  const int ie = min(ib + ta.cx, ta.hi[0]);
  const long long N = ta.N;
  const long long sx = (long long)ta.Ny * ta.Nz, sy = ta.Nz;
  const int kc = min(k, ta.hi[2] - 1), jc = min(j, ta.hi[1] - 1);
  const long long col = (long long)jc * sy + kc;
  const unsigned colb = (unsigned)(col * (long long)sizeof(R));  // < one plane in bytes
  const unsigned psb = (unsigned)(sx * (long long)sizeof(R));    // plane size in bytes
  // field q of plane i: scalar byte offset i*psb within the field
  // df/f traffic a stage does not need goes through a descriptor with no
  // records: out-of-range buffer loads return 0 and stores are dropped,
  // without a memory request, and the instruction stream stays the same
  typedef typename FieldsOf<R, WIDE>::type FA;
  const FA rf(f, N, 9, true), rd(df, N, 9, ta.dfmode & DF_LOAD), rds(df, N, 9, ta.dfmode & DF_STORE),
      rn(fn, N, 9, ta.dfmode & F_STORE);
  const bool yin = j >= ta.mc[1] && j < ta.mrb[1];
  const bool zin = k >= ta.mc[2] && k < ta.mrb[2];
  // block constants in the compute precision (uniform, scalar registers)
  C kcv[3], kcg2[3], kcl[3], kcg[3], kA, kB;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if constexpr (sizeof(C) == 4) {
      kcv[d] = ta.fcv[d]; kcg2[d] = ta.fcg2[d]; kcl[d] = ta.fcl[d]; kcg[d] = ta.fcg[d];
    } else {
      kcv[d] = ta.cv[d]; kcg2[d] = ta.cg2[d]; kcl[d] = ta.cl[d]; kcg[d] = ta.cg[d];
    }
  }
  if constexpr (sizeof(C) == 4) { kA = ta.fA; kB = ta.fB; } else { kA = ta.A; kB = ta.B; }
  C cd[r + 1];
#pragma unroll
  for (int n = 0; n <= r; n++) cd[n] = (C)c_fd[0][r + n];
  constexpr int NZH = TY * 2 * r, NH = NZH + 2 * r * TZ;
  auto xplane = [&](int i) { return (unsigned)min(max(i, ta.lo[0]), ta.hi[0] - 1) * psb; };
  // x window: fields 0..5 = vx vy vz sxx sxy sxz at planes i-r .. i+r (clamped)
  R win[6][W];
#pragma unroll
  for (int n = 0; n < W - 1; n++) {
    const unsigned pw = xplane(ib - r + n);
#pragma unroll
    for (int q = 0; q < 6; q++) win[q][n] = rf.ld(q, colb, pw);
  }
  constexpr int tf[8] = {0, 1, 2, 4, 5, 6, 7, 8};  // f index of each tile field
  constexpr int yq[6] = {0, 1, 2, 4, 6, 7};  // vx vy vz sxy syy syz
  constexpr int zq[6] = {0, 1, 2, 5, 7, 8};  // vx vy vz sxz syz szz
  // halo entries this thread loads each plane (cross-shaped ring of the tile)
  constexpr int NHT = (NH + NT - 1) / NT;
  int hly[NHT], hlz[NHT];
  unsigned hoff[NHT];  // byte offset of the halo entry within a plane
#pragma unroll
  for (int u = 0; u < NHT; u++) {
    const int e = threadIdx.x + u * NT;
    int ly = 0, lz = 0;
    if (e < NZH) {
      ly = r + e / (2 * r);
      const int h = e % (2 * r);
      lz = (h < r) ? h : TZ + h;
    } else if (e < NH) {
      const int e2 = e - NZH;
      const int h = e2 / TZ;
      ly = (h < r) ? h : TY + h;
      lz = r + e2 % TZ;
    }
    hly[u] = (e < NH) ? ly : -1;
    hlz[u] = lz;
    const int gj = min(max(jb - r + ly, ta.lo[1]), ta.hi[1] - 1);
    const int gk = min(max(kb - r + lz, ta.lo[2]), ta.hi[2] - 1);
    hoff[u] = (unsigned)(((long long)gj * sy + gk) * (long long)sizeof(R));
  }
  // all loads of one plane are issued together, one plane ahead of its use
  R nwin[6], nown[3], nhal[NHT][8];
  auto issue = [&](int ip) {
    const unsigned pw = xplane(ip + r), pp = (unsigned)ip * psb;
#pragma unroll
    for (int q = 0; q < 6; q++) nwin[q] = rf.ld(q, colb, pw);
#pragma unroll
    for (int q = 0; q < 3; q++) nown[q] = rf.ld(6 + q, colb, pp);
#pragma unroll
    for (int u = 0; u < NHT; u++)
      if (hly[u] >= 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) nhal[u][q] = rf.ld(tf[q], hoff[u], pp);
      }
  };
  issue(ib);
  for (int i = ib; i < ie; i++) {
    const long long pl = (long long)i * sx;
    const unsigned plb = (unsigned)i * psb;
#pragma unroll
    for (int q = 0; q < 6; q++) win[q][W - 1] = nwin[q];
    R own[3];
#pragma unroll
    for (int q = 0; q < 3; q++) own[q] = nown[q];
    tile[0][ty + r][tz + r] = win[0][r];
    tile[1][ty + r][tz + r] = win[1][r];
    tile[2][ty + r][tz + r] = win[2][r];
    tile[3][ty + r][tz + r] = win[4][r];
    tile[4][ty + r][tz + r] = win[5][r];
    tile[5][ty + r][tz + r] = own[0];
    tile[6][ty + r][tz + r] = own[1];
    tile[7][ty + r][tz + r] = own[2];
#pragma unroll
    for (int u = 0; u < NHT; u++)
      if (hly[u] >= 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) tile[q][hly[u]][hlz[u]] = nhal[u][q];
      }
    lds_barrier();
    if (i + 1 < ie) issue(i + 1);
    // df of this plane is first needed at the end of the update: loading it
    // here rather than a plane ahead keeps it out of the loop-carried state
    R odf[9];
#pragma unroll
    for (int q = 0; q < 9; q++) odf[q] = rd.ld(q, colb, plb);
    if (active) {
      C dxv[6], dyv[6], dzv[6];
      const int jw = __builtin_amdgcn_readfirstlane(j);  // a wave is one y row
      if (i >= ta.mc[0] && i < ta.mrb[0]) {
#pragma unroll
        for (int q = 0; q < 6; q++)  // d/dx of vx vy vz sxx sxy sxz
          dxv[q] = central<C, r>(cd, [&](int n) { return (C)win[q][n]; });
      } else {
        edge_x<R, C, S>(dxv, rf, psb, colb, i, ta.lo[0], ta.hi[0]);
      }
      constexpr int ymap[6] = {0, 1, 2, 3, 5, 6}, zmap[6] = {0, 1, 2, 4, 6, 7};
      if (yin) {
#pragma unroll
        for (int q = 0; q < 6; q++)
          dyv[q] = central<C, r>(cd, [&](int n) { return (C)tile[ymap[q]][ty + n][tz + r]; });
      } else {
        // one-sided row: taps from the LDS tile when they lie inside it
        const bool left = (j - ta.lo[1]) < 2 * (S - 1);
        const int row = left ? j - ta.lo[1] + 1 : ta.hi[1] - j;
        const int t0 = left ? ta.lo[1] : ta.hi[1] - NTB;  // lowest tap
        if (t0 >= jb - r && t0 + NTB <= jb + TY + r) {
#pragma unroll
          for (int q = 0; q < 6; q++) {
            C sy_ = 0.;
#pragma unroll
            for (int n = 0; n < NTB; n++) {
              const int tj = left ? ta.lo[1] + n : ta.hi[1] - 1 - n;
              sy_ += sfd[row - 1][n] * (C)tile[ymap[q]][tj - (jb - r)][tz + r];
            }
            dyv[q] = left ? sy_ : -sy_;
          }
        } else {
          edge_plane<R, C, S>(dyv, rf, plb, yq, k, (int)sy, jw, ta.lo[1], ta.hi[1], sfd);
        }
      }
      if (zin) {
#pragma unroll
        for (int q = 0; q < 6; q++)
          dzv[q] = central<C, r>(cd, [&](int n) { return (C)tile[zmap[q]][ty + r][tz + n]; });
      } else {
        const bool left = (k - ta.lo[2]) < 2 * (S - 1);
        const int row = left ? k - ta.lo[2] + 1 : ta.hi[2] - k;
        const int t0 = left ? ta.lo[2] : ta.hi[2] - NTB;
        if (t0 >= kb - r && t0 + NTB <= kb + TZ + r) {
#pragma unroll
          for (int q = 0; q < 6; q++) {
            C sz_ = 0.;
#pragma unroll
            for (int n = 0; n < NTB; n++) {
              const int tk = left ? ta.lo[2] + n : ta.hi[2] - 1 - n;
              sz_ += sfd[row - 1][n] * (C)tile[zmap[q]][ty + r][tk - (kb - r)];
            }
            dzv[q] = left ? sz_ : -sz_;
          }
        } else {
          edge_plane<R, C, S>(dzv, rf, plb, zq, jw * (int)sy, 1, k, ta.lo[2], ta.hi[2], sfd);
        }
      }
      const long long id = pl + (long long)j * sy + k;
      C acc[9];
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q] = (C)odf[q] * kA;
      acc[0] += kcv[0] * dxv[3] + kcv[1] * dyv[3] + kcv[2] * dzv[3];
      acc[1] += kcv[0] * dxv[4] + kcv[1] * dyv[4] + kcv[2] * dzv[4];
      acc[2] += kcv[0] * dxv[5] + kcv[1] * dyv[5] + kcv[2] * dzv[5];
      acc[3] += kcg2[0] * dxv[0] + kcl[1] * dyv[1] + kcl[2] * dzv[2];
      acc[4] += kcg[0] * dxv[1] + kcg[1] * dyv[0];
      acc[5] += kcg[0] * dxv[2] + kcg[2] * dzv[0];
      acc[6] += kcl[0] * dxv[0] + kcg2[1] * dyv[1] + kcl[2] * dzv[2];
      acc[7] += kcg[1] * dyv[2] + kcg[2] * dzv[1];
      acc[8] += kcl[0] * dxv[0] + kcl[1] * dyv[1] + kcg2[2] * dzv[2];
      C fo[9];
#pragma unroll
      for (int q = 0; q < 6; q++) fo[q] = (C)win[q][r];
      fo[6] = (C)own[0];
      fo[7] = (C)own[1];
      fo[8] = (C)own[2];
      // external SAT: on a Cartesian face with homogeneous material the
      // boundary.cpp:265-492 penalty is a fixed linear map of the point's fields
      const int c3[3] = {i, j, k};
#pragma unroll
      for (int e = 0; e < 6; e++) {
        if (!ta.face_active[e]) continue;
        constexpr int sid[3][3] = {{3, 4, 5}, {4, 6, 7}, {5, 7, 8}};  // stress index of (row, col)
        const int d = e >> 1, a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
        if (c3[d] != ((e & 1) ? ta.hi[d] - 1 : ta.lo[d])) continue;
        // read in the loop from device memory: kernel arguments would be
        // hoisted out of the loop as 96 live registers
        const C* m = reinterpret_cast<const C*>(ta.fcoef) + e * 16;
        const C vd = fo[d], sdd = fo[sid[d][d]];
        acc[d] += (C)m[0] * vd + (C)m[1] * sdd;
        acc[sid[d][d]] += (C)m[2] * vd + (C)m[3] * sdd;
        acc[sid[a][a]] += (C)m[4] * vd + (C)m[5] * sdd;
        acc[sid[b][b]] += (C)m[6] * vd + (C)m[7] * sdd;
        const C va = fo[a], sda = fo[sid[d][a]], vb = fo[b], sdb = fo[sid[d][b]];
        acc[a] += (C)m[8] * va + (C)m[9] * sda;
        acc[sid[d][a]] += (C)m[10] * va + (C)m[11] * sda;
        acc[b] += (C)m[12] * vb + (C)m[13] * sdb;
        acc[sid[d][b]] += (C)m[14] * vb + (C)m[15] * sdb;
      }
      // interface SAT terms written by fault_kernel (interface.cpp:826-962)
      for (int e = 0; e < ta.nif; e++) {
        if (c3[ta.if_dir[e]] != ta.if_fixed[e]) continue;
        const unsigned o = (unsigned)(((c3[ta.if_a[e]] - ta.if_orga[e]) * ta.if_n1[e] + (c3[ta.if_b[e]] - ta.if_orgb[e])) * 8);
        const long long n = ta.if_n[e];
        const double* sn = ta.if_sat[e] + (long long)(ta.if_side[e] * 2 + 0) * 9 * n;
        const double* sv = sn + 9 * n;
#pragma unroll
        for (int q = 0; q < 9; q++) acc[q] -= (C)(sn[q * n + o / 8] + sv[q * n + o / 8]);
      }
#pragma unroll
      for (int q = 0; q < 9; q++) {
        rds.st(q, (R)acc[q], colb, plb);
        C v = fo[q] + kB * acc[q];
        if (q >= 3 && ta.roundtrip) {  // set_stress / remove_stress rounding (FP64 only)
          double w = (double)v;
          w += ta.s0[q - 3];
          if (ta.hetstress) w += hs[(q - 3) * N + id];
          w -= ta.s0[q - 3];
          if (ta.hetstress) w -= hs[(q - 3) * N + id];
          v = (C)w;
        }
        rn.st(q, (R)v, colb, plb);
      }
    }
    lds_barrier();  // every wave is done reading this plane's tile
    // slide the window
#pragma unroll
    for (int n = 0; n < W - 1; n++)
#pragma unroll
      for (int q = 0; q < 6; q++) win[q][n] = win[q][n + 1];
  }
}

// ---------------------------------------------------------------------------
// Rectilinear 3-D block kernel with row-shaped tiles.
//
// Same update as bulk_tiled3d, but a workgroup owns TY whole rows of the
// block (all nz points of each), threads mapped linearly over the TY*nz
// points, so the memory it touches per plane and field is one contiguous run
// of TY*nz values: whole DRAM pages instead of 256-byte pieces of them.  The
// LDS tile [8][TY + 2r][nz] has y halos only (a row is complete in z), so
// every one-sided z row reads its taps from LDS; the z derivative is
// branch-free (per-lane coefficient row and tap offset) because rows do not
// align with 64-lane waves.  One-sided y rows read LDS when their taps lie in
// the tile, else global memory.  NHT halo entries per thread per plane;
// rows of at most 256 points.
//
// NBUF = 2 double-buffers the LDS tile: plane i's tile is written while
// plane i-1's buffer may still be read, so one barrier per plane suffices
// (a wave writing buffer b at plane i+1 has passed plane i's barrier, which
// every wave reaches only after its plane i-1 reads of b).
template <typename R, typename C, int S, int TY, int NTMAX, int NHT, int WPE, int NBUF = 1, int ZS = 256,
          bool WIDE = false>
__global__ void __launch_bounds__(NTMAX) __attribute__((amdgpu_waves_per_eu(WPE)))
    bulk_rows3d(Tiled3D ta, const R* __restrict__ f, R* __restrict__ fn, R* __restrict__ df,
                const double* __restrict__ hs) {
#pragma clang fp contract(fast)
  constexpr int r = S - 1;
  constexpr int W = 2 * r + 1;
  constexpr int LY = TY + 2 * r;
  constexpr int NRB = 2 * (S - 1), NTB = 3 * (S - 1);
  // LDS rows padded to ZS >= nz entries (a compile-time stride, so every
  // tap is a lane base plus an immediate offset)
  extern __shared__ char smem_raw[];
  R* const tbase = reinterpret_cast<R*>(smem_raw);  // [NBUF][8][LY][ZS]
  // z derivative coefficients: row 0 interior (taps -r..r, zero padded),
  // rows 1..NRB left one-sided rows, NRB+1..2NRB right rows (forward order)
  __shared__ C szc[1 + 2 * NRB][NTB];
  __shared__ C sfd[NRB][NTB];
  const int nz = ta.hi[2] - ta.lo[2];
  const int NT = blockDim.x;
  const int t = threadIdx.x;
  for (int e = t; e < (1 + 2 * NRB) * NTB; e += NT) {
    const int rw = e / NTB, n = e % NTB;
    C c = 0;
    if (rw == 0) c = n < W ? (C)c_fd[0][n] : (C)0;
    else if (rw <= NRB) c = (C)c_fd[rw][n];
    else c = -(C)c_fd[rw - NRB][NTB - 1 - n];
    szc[rw][n] = c;
  }
  if (t < NRB * NTB) sfd[t / NTB][t % NTB] = (C)c_fd[1 + t / NTB][t % NTB];
  const int nxcd = 8;
  const int per = (ta.ntiles + nxcd - 1) / nxcd;
  const int tile_id = (blockIdx.x % nxcd) * per + blockIdx.x / nxcd;
  if (tile_id >= ta.ntiles) return;
  // Tile order: strips of grp_x x chunks spanning every y tile, cut into
  // patches of grp_y y tiles (the last strip and patch may be narrower), y
  // fastest inside a patch.  The workgroups an XCD runs at one time then
  // cover a compact patch, so the y-halo rows and the x-window planes that
  // neighbouring tiles share are read while the first reader's copy is still
  // in that XCD's L2.  grp_x = 1, grp_y >= nyt is the plain y-fastest order.
  int yt, xc;
  {
    const int gx = ta.grp_x, gy = ta.grp_y;
    const int sx = tile_id / (ta.nyt * gx), rem = tile_id - sx * ta.nyt * gx;
    const int wx = min(gx, ta.nxc - sx * gx);
    const int by = rem / (gy * wx), h = min(gy, ta.nyt - by * gy);
    const int w = rem - by * gy * wx;
    yt = by * gy + w % h;
    xc = sx * gx + w / h;
  }
  const int ty = t / nz, kk = t - ty * nz;
  const int jb = ta.lo[1] + yt * TY, j = jb + ty, k = ta.lo[2] + kk;
  const bool active = ty < TY && j < ta.hi[1];
  const int ib = ta.lo[0] + xc * ta.cx;
  const int ie = min(ib + ta.cx, ta.hi[0]);
  const long long N = ta.N;
  const long long sx = (long long)ta.Ny * ta.Nz, sy = ta.Nz;
  const int jc = min(j, ta.hi[1] - 1), tyc = min(ty, TY - 1);
  const unsigned colb = (unsigned)(((long long)jc * sy + k) * (long long)sizeof(R));
  const unsigned psb = (unsigned)(sx * (long long)sizeof(R));
  // df/f traffic a stage does not need goes through a descriptor with no
  // records: out-of-range buffer loads return 0 and stores are dropped,
  // without a memory request, and the instruction stream stays the same
  typedef typename FieldsOf<R, WIDE>::type FA;
  const FA rf(f, N, 9, true), rd(df, N, 9, ta.dfmode & DF_LOAD), rds(df, N, 9, ta.dfmode & DF_STORE),
      rn(fn, N, 9, ta.dfmode & F_STORE);
  const bool yin = j >= ta.mc[1] && j < ta.mrb[1];
  // z row of this lane: coefficient row and first tap column
  int zrow = 0, zoff = -r;
  if (k < ta.mc[2]) { zrow = kk + 1; zoff = -kk; }
  else if (k >= ta.mrb[2]) { zrow = NRB + (nz - kk); zoff = nz - NTB - kk; }
  C kcv[3], kcg2[3], kcl[3], kcg[3], kA, kB;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if constexpr (sizeof(C) == 4) {
      kcv[d] = ta.fcv[d]; kcg2[d] = ta.fcg2[d]; kcl[d] = ta.fcl[d]; kcg[d] = ta.fcg[d];
    } else {
      kcv[d] = ta.cv[d]; kcg2[d] = ta.cg2[d]; kcl[d] = ta.cl[d]; kcg[d] = ta.cg[d];
    }
  }
  if constexpr (sizeof(C) == 4) { kA = ta.fA; kB = ta.fB; } else { kA = ta.A; kB = ta.B; }
  C cd[r + 1];
#pragma unroll
  for (int n = 0; n <= r; n++) cd[n] = (C)c_fd[0][r + n];
  auto xplane = [&](int i) { return (unsigned)min(max(i, ta.lo[0]), ta.hi[0] - 1) * psb; };
  R win[6][W];
#pragma unroll
  for (int n = 0; n < W - 1; n++) {
    const unsigned pw = xplane(ib - r + n);
#pragma unroll
    for (int q = 0; q < 6; q++) win[q][n] = rf.ld(q, colb, pw);
  }
  constexpr int tf[8] = {0, 1, 2, 4, 5, 6, 7, 8};  // f index of each tile field
  constexpr int yq[6] = {0, 1, 2, 4, 6, 7};        // vx vy vz sxy syy syz
  // halo entries: rows 0..r-1 and TY+r..TY+2r-1 of the tile, all nz columns;
  // only the six fields with a y derivative (tile fields hq) are read there
  constexpr int hq[6] = {0, 1, 2, 3, 5, 6};
  int hrow[NHT];
  unsigned hoff[NHT];
#pragma unroll
  for (int u = 0; u < NHT; u++) {
    const int e = t + u * NT;
    const int hr = e / nz, hk = e - hr * nz;
    const int lr = hr < r ? hr : TY + hr;
    hrow[u] = hr < 2 * r ? lr * ZS + hk : -1;
    const int gj = min(max(jb - r + lr, ta.lo[1]), ta.hi[1] - 1);
    hoff[u] = (unsigned)(((long long)gj * sy + ta.lo[2] + (hr < 2 * r ? hk : 0)) * (long long)sizeof(R));
  }
  R nwin[6], nown[3], nhal[NHT][6];
  auto issue = [&](int ip) {
    const unsigned pw = xplane(ip + r), pp = (unsigned)ip * psb;
#pragma unroll
    for (int q = 0; q < 6; q++) nwin[q] = rf.ld(q, colb, pw);
#pragma unroll
    for (int q = 0; q < 3; q++) nown[q] = rf.ld(6 + q, colb, pp);
#pragma unroll
    for (int u = 0; u < NHT; u++)
      if (hrow[u] >= 0) {
#pragma unroll
        for (int q = 0; q < 6; q++) nhal[u][q] = rf.ld(tf[hq[q]], hoff[u], pp);
      }
  };
  constexpr int LPL = LY * ZS;  // tile entries per field
  const int own_at = (tyc + r) * ZS + kk;
  issue(ib);
  for (int i = ib; i < ie; i++) {
    const long long pl = (long long)i * sx;
    const unsigned plb = (unsigned)i * psb;
    R* const tile = tbase + (NBUF == 2 ? ((i - ib) & 1) * (8 * LPL) : 0);
#pragma unroll
    for (int q = 0; q < 6; q++) win[q][W - 1] = nwin[q];
    R own[3];
#pragma unroll
    for (int q = 0; q < 3; q++) own[q] = nown[q];
    if (ty < TY) {
      tile[0 * LPL + own_at] = win[0][r];
      tile[1 * LPL + own_at] = win[1][r];
      tile[2 * LPL + own_at] = win[2][r];
      tile[3 * LPL + own_at] = win[4][r];
      tile[4 * LPL + own_at] = win[5][r];
      tile[5 * LPL + own_at] = own[0];
      tile[6 * LPL + own_at] = own[1];
      tile[7 * LPL + own_at] = own[2];
    }
#pragma unroll
    for (int u = 0; u < NHT; u++)
      if (hrow[u] >= 0) {
#pragma unroll
        for (int q = 0; q < 6; q++) tile[hq[q] * LPL + hrow[u]] = nhal[u][q];
      }
    lds_barrier();
    // every wave issues the next plane's loads and this plane's df loads
    // before any wave starts the arithmetic (measured: 2.5 % faster FP32)
    __builtin_amdgcn_s_setprio(3);
    R odf[9];
    if constexpr (sizeof(R) == 4) {
      if (i + 1 < ie) issue(i + 1);
#pragma unroll
      for (int q = 0; q < 9; q++) odf[q] = rd.ld(q, colb, plb);
    } else {
      // FP64: this plane's df first, so that waiting for it does not wait
      // for the next plane's loads too (vmcnt counts in issue order;
      // measured 1.243 -> 1.232 ms per block launch, FP32 unchanged)
#pragma unroll
      for (int q = 0; q < 9; q++) odf[q] = rd.ld(q, colb, plb);
      asm volatile("" ::: "memory");
      if (i + 1 < ie) issue(i + 1);
    }
    __builtin_amdgcn_s_setprio(0);
    if (active) {
      C dxv[6], dyv[6], dzv[6];
      if (i >= ta.mc[0] && i < ta.mrb[0]) {
#pragma unroll
        for (int q = 0; q < 6; q++)
          dxv[q] = central<C, r>(cd, [&](int n) { return (C)win[q][n]; });
      } else {
        edge_x<R, C, S>(dxv, rf, psb, colb, i, ta.lo[0], ta.hi[0]);
      }
      constexpr int ymap[6] = {0, 1, 2, 3, 5, 6}, zmap[6] = {0, 1, 2, 4, 6, 7};
      if (yin) {
#pragma unroll
        for (int q = 0; q < 6; q++)
          dyv[q] = central<C, r>(cd, [&](int n) { return (C)tile[ymap[q] * LPL + (ty + n) * ZS + kk]; });
      } else {
        const bool left = (j - ta.lo[1]) < NRB;
        const int row = left ? j - ta.lo[1] + 1 : ta.hi[1] - j;
        const int t0 = left ? ta.lo[1] : ta.hi[1] - NTB;
        if (t0 >= jb - r && t0 + NTB <= jb + TY + r) {
#pragma unroll
          for (int q = 0; q < 6; q++) {
            C sy_ = 0.;
#pragma unroll
            for (int n = 0; n < NTB; n++) {
              const int tj = left ? ta.lo[1] + n : ta.hi[1] - 1 - n;
              sy_ += sfd[row - 1][n] * (C)tile[ymap[q] * LPL + (tj - (jb - r)) * ZS + kk];
            }
            dyv[q] = left ? sy_ : -sy_;
          }
        } else {
          edge_plane<R, C, S>(dyv, rf, plb, yq, k, (int)sy, j, ta.lo[1], ta.hi[1], sfd);
        }
      }
      {  // z: NTB taps from column kk + zoff with the lane's coefficient row
        const int zb = (ty + r) * ZS + kk + zoff;
#pragma unroll
        for (int q = 0; q < 6; q++) {
          C sz = szc[zrow][0] * (C)tile[zmap[q] * LPL + zb];
#pragma unroll
          for (int n = 1; n < NTB; n++) sz += szc[zrow][n] * (C)tile[zmap[q] * LPL + zb + n];
          dzv[q] = sz;
        }
      }
      const long long id = pl + (long long)j * sy + k;
      C acc[9];
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q] = (C)odf[q] * kA;
      acc[0] += kcv[0] * dxv[3] + kcv[1] * dyv[3] + kcv[2] * dzv[3];
      acc[1] += kcv[0] * dxv[4] + kcv[1] * dyv[4] + kcv[2] * dzv[4];
      acc[2] += kcv[0] * dxv[5] + kcv[1] * dyv[5] + kcv[2] * dzv[5];
      acc[3] += kcg2[0] * dxv[0] + kcl[1] * dyv[1] + kcl[2] * dzv[2];
      acc[4] += kcg[0] * dxv[1] + kcg[1] * dyv[0];
      acc[5] += kcg[0] * dxv[2] + kcg[2] * dzv[0];
      acc[6] += kcl[0] * dxv[0] + kcg2[1] * dyv[1] + kcl[2] * dzv[2];
      acc[7] += kcg[1] * dyv[2] + kcg[2] * dzv[1];
      acc[8] += kcl[0] * dxv[0] + kcl[1] * dyv[1] + kcg2[2] * dzv[2];
      C fo[9];
#pragma unroll
      for (int q = 0; q < 6; q++) fo[q] = (C)win[q][r];
      fo[6] = (C)own[0];
      fo[7] = (C)own[1];
      fo[8] = (C)own[2];
      const int c3[3] = {i, j, k};
#pragma unroll
      for (int e = 0; e < 6; e++) {
        if (!ta.face_active[e]) continue;
        constexpr int sid[3][3] = {{3, 4, 5}, {4, 6, 7}, {5, 7, 8}};
        const int d = e >> 1, a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
        if (c3[d] != ((e & 1) ? ta.hi[d] - 1 : ta.lo[d])) continue;
        const C* m = reinterpret_cast<const C*>(ta.fcoef) + e * 16;
        const C vd = fo[d], sdd = fo[sid[d][d]];
        acc[d] += (C)m[0] * vd + (C)m[1] * sdd;
        acc[sid[d][d]] += (C)m[2] * vd + (C)m[3] * sdd;
        acc[sid[a][a]] += (C)m[4] * vd + (C)m[5] * sdd;
        acc[sid[b][b]] += (C)m[6] * vd + (C)m[7] * sdd;
        const C va = fo[a], sda = fo[sid[d][a]], vb = fo[b], sdb = fo[sid[d][b]];
        acc[a] += (C)m[8] * va + (C)m[9] * sda;
        acc[sid[d][a]] += (C)m[10] * va + (C)m[11] * sda;
        acc[b] += (C)m[12] * vb + (C)m[13] * sdb;
        acc[sid[d][b]] += (C)m[14] * vb + (C)m[15] * sdb;
      }
      for (int e = 0; e < ta.nif; e++) {
        if (c3[ta.if_dir[e]] != ta.if_fixed[e]) continue;
        const long long o = (long long)(c3[ta.if_a[e]] - ta.if_orga[e]) * ta.if_n1[e] + (c3[ta.if_b[e]] - ta.if_orgb[e]);
        const long long n = ta.if_n[e];
        const double* sn = ta.if_sat[e] + (long long)(ta.if_side[e] * 2 + 0) * 9 * n;
        const double* sv = sn + 9 * n;
#pragma unroll
        for (int q = 0; q < 9; q++) acc[q] -= (C)(sn[q * n + o] + sv[q * n + o]);
      }
#pragma unroll
      for (int q = 0; q < 9; q++) {
        rds.st(q, (R)acc[q], colb, plb);
        C v = fo[q] + kB * acc[q];
        if (q >= 3 && ta.roundtrip) {
          double w = (double)v;
          w += ta.s0[q - 3];
          if (ta.hetstress) w += hs[(q - 3) * N + id];
          w -= ta.s0[q - 3];
          if (ta.hetstress) w -= hs[(q - 3) * N + id];
          v = (C)w;
        }
        rn.st(q, (R)v, colb, plb);
      }
    }
    if constexpr (NBUF == 1) lds_barrier();  // every wave is done reading this plane's tile
#pragma unroll
    for (int n = 0; n < W - 1; n++)
#pragma unroll
      for (int q = 0; q < 6; q++) win[q][n] = win[q][n + 1];
  }
}

// ---------------------------------------------------------------------------
// Rectilinear 2-D mode II block kernel (fields vx vy sxx sxy syy).
//
// On a Cartesian block the update of block.cpp:1048-1278 reduces to
// per-direction derivatives scaled by block constants.  A workgroup of TW
// threads owns TW consecutive y points (y is the contiguous axis in 2-D) and
// marches along x over a chunk of planes:
//   * x derivatives of vx vy sxx sxy from a register window over 2r+1 planes;
//   * y derivatives of vx vy sxy syy from an LDS row with a halo of 3(s-1)
//     points on each side, wide enough that every one-sided SBP row of the
//     block reads its taps from LDS;
//   * external SAT as the face's 10 non-zero linear-map entries, interface
//     SAT from the fault buffer, df = A*df + ..., f_new = f + B*df.
// The per-tile routine is shared by the single-block launch and the batched
// (ensemble) launch, so both give bit-identical results for one instance.
struct Tiled2D {
  int lo[2], hi[2];   // block box (x, y), global indices
  int mc[2], mrb[2];  // interior rows [mc, mrb) per direction
  int Ny;             // y size of the grid (row stride)
  long long N;        // grid points
  int cx;             // x planes per chunk
  int tw;             // y points per workgroup (the kernel's TW)
  int nyt, nxc, ntiles;
  double cv[2], cg2[2], cl[2], cg[2], A, B;
  float fcv[2], fcg2[2], fcl[2], fcg[2], fA, fB;
  int roundtrip, hetstress;
  int dfmode = DF_ALL;  // DF_LOAD / DF_STORE / F_STORE bits (single-block launch)
  int wide;           // 1: one buffer descriptor per field (nf N elements above 4 GiB)
  double s0[3];       // sxx sxy syy
  int face_active[4];
  const void* fcoef;  // [4 faces][10] external SAT entries in the bulk precision
  const void* smap[4];  // curvilinear blocks: per face, per point SAT map [n][rows][5] (bulk precision)
  int nif;
  int if_dir[4], if_fixed[4], if_a[4], if_orga[4], if_n[4], if_side[4];
  const double* if_sat[4];
  // arrays of this grid (the batched launch gives each instance its own)
  const void* f[2];   // f, indexed by the parity of the stage
  void* df;
  const double* hs;
};

template <typename R, typename C, int S, int TW, bool WIDE>
__device__ __forceinline__ void tile2d(const Tiled2D& ta, int tile_id, int par, int dfm) {
#pragma clang fp contract(fast)
  constexpr int r = S - 1, W = 2 * r + 1;
  constexpr int NRB = 2 * (S - 1), NTB = 3 * (S - 1);
  constexpr int H = NTB;            // LDS halo per side
  constexpr int LW = TW + 2 * H;
  // double-buffered by plane parity: plane i writes one buffer while a
  // slower wave may still read plane i-1's, so one barrier per plane suffices
  __shared__ R row[2][4][LW];       // 0 vx 1 vy 2 sxy 3 syy
  __shared__ C sfd[NRB][NTB];
  const R* __restrict__ f = reinterpret_cast<const R*>(ta.f[par]);
  R* __restrict__ fn = const_cast<R*>(reinterpret_cast<const R*>(ta.f[par ^ 1]));
  R* __restrict__ df = reinterpret_cast<R*>(ta.df);
  const int t = threadIdx.x;
  if (t < NRB * NTB) sfd[t / NTB][t % NTB] = (C)c_fd[1 + t / NTB][t % NTB];
  const int yt = tile_id % ta.nyt, xc = tile_id / ta.nyt;
  const int jb = ta.lo[1] + yt * TW, j = jb + t;
  const bool active = j < ta.hi[1];
  const int ib = ta.lo[0] + xc * ta.cx, ie = min(ib + ta.cx, ta.hi[0]);
  const long long N = ta.N;
  const int jc = min(j, ta.hi[1] - 1);
  const unsigned colb = (unsigned)jc * (unsigned)sizeof(R);
  const unsigned psb = (unsigned)ta.Ny * (unsigned)sizeof(R);
  // df/f traffic a stage does not need: descriptors with no records (loads
  // return 0, stores are dropped, no memory request)
  typedef typename FieldsOf<R, WIDE>::type FA;
  const FA rf(f, N, 5, true), rd(df, N, 5, dfm & DF_LOAD), rds(df, N, 5, dfm & DF_STORE), rn(fn, N, 5, dfm & F_STORE);
  const bool yin = j >= ta.mc[1] && j < ta.mrb[1];
  C kcv[2], kcg2[2], kcl[2], kcg[2], kA, kB;
#pragma unroll
  for (int d = 0; d < 2; d++) {
    if constexpr (sizeof(C) == 4) {
      kcv[d] = ta.fcv[d]; kcg2[d] = ta.fcg2[d]; kcl[d] = ta.fcl[d]; kcg[d] = ta.fcg[d];
    } else {
      kcv[d] = ta.cv[d]; kcg2[d] = ta.cg2[d]; kcl[d] = ta.cl[d]; kcg[d] = ta.cg[d];
    }
  }
  if constexpr (sizeof(C) == 4) { kA = ta.fA; kB = ta.fB; } else { kA = ta.A; kB = ta.B; }
  C cd[r + 1];
#pragma unroll
  for (int n = 0; n <= r; n++) cd[n] = (C)c_fd[0][r + n];
  auto xplane = [&](int i) { return (unsigned)min(max(i, ta.lo[0]), ta.hi[0] - 1) * psb; };
  // halo entry of this thread (threads 0..2H-1): row position and y index
  const bool hal = t < 2 * H;
  const int hpos = t < H ? t : TW + t;  // position in the LDS row
  const int hj = min(max(jb - H + hpos, ta.lo[1]), ta.hi[1] - 1);
  const unsigned hoff = (unsigned)hj * (unsigned)sizeof(R);
  // x window: vx vy sxx sxy at planes i-r .. i+r (clamped)
  R win[4][W];
#pragma unroll
  for (int n = 0; n < W - 1; n++) {
    const unsigned pw = xplane(ib - r + n);
#pragma unroll
    for (int q = 0; q < 4; q++) win[q][n] = rf.ld(q, colb, pw);
  }
  constexpr int tf[4] = {0, 1, 3, 4};  // f index of each LDS row field
  R nwin[4], nown, nhal[4];
  auto issue = [&](int ip) {
    const unsigned pw = xplane(ip + r), pp = (unsigned)ip * psb;
#pragma unroll
    for (int q = 0; q < 4; q++) nwin[q] = rf.ld(q, colb, pw);
    nown = rf.ld(4, colb, pp);
    if (hal) {
#pragma unroll
      for (int q = 0; q < 4; q++) nhal[q] = rf.ld(tf[q], hoff, pp);
    }
  };
  issue(ib);
  for (int i = ib; i < ie; i++) {
    auto* const rb = row[(i - ib) & 1];  // this plane's LDS row buffer
    const unsigned plb = (unsigned)i * psb;
#pragma unroll
    for (int q = 0; q < 4; q++) win[q][W - 1] = nwin[q];
    const R own = nown;
    rb[0][H + t] = win[0][r];
    rb[1][H + t] = win[1][r];
    rb[2][H + t] = win[3][r];
    rb[3][H + t] = own;
    if (hal) {
#pragma unroll
      for (int q = 0; q < 4; q++) rb[q][hpos] = nhal[q];
    }
    lds_barrier();
    if (i + 1 < ie) issue(i + 1);
    R odf[5];
#pragma unroll
    for (int q = 0; q < 5; q++) odf[q] = rd.ld(q, colb, plb);
    if (active) {
      C dxv[4], dyv[4];  // d/dx of vx vy sxx sxy; d/dy of vx vy sxy syy
      if (i >= ta.mc[0] && i < ta.mrb[0]) {
#pragma unroll
        for (int q = 0; q < 4; q++) dxv[q] = central<C, r>(cd, [&](int n) { return (C)win[q][n]; });
      } else {  // one-sided row along x: taps are whole planes (wave-uniform)
        const bool left = i - ta.lo[0] < NRB;
        const int rw = left ? i - ta.lo[0] + 1 : ta.hi[0] - i;
#pragma unroll
        for (int n = 0; n < NTB; n++) {
          const int tap = left ? ta.lo[0] + n : ta.hi[0] - 1 - n;
          const C c = (C)c_fd[rw][n];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const C v = (C)rf.ld(q, colb, (unsigned)tap * psb);
            dxv[q] = n == 0 ? c * v : dxv[q] + c * v;
          }
        }
        if (!left)
#pragma unroll
          for (int q = 0; q < 4; q++) dxv[q] = -dxv[q];
      }
      if (yin) {
#pragma unroll
        for (int q = 0; q < 4; q++) dyv[q] = central<C, r>(cd, [&](int n) { return (C)rb[q][t + H - r + n]; });
      } else {  // one-sided row along y: taps in the LDS row
        const bool left = j - ta.lo[1] < NRB;
        const int rw = left ? j - ta.lo[1] + 1 : ta.hi[1] - j;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          C sy = 0.;
#pragma unroll
          for (int n = 0; n < NTB; n++) {
            const int tj = left ? ta.lo[1] + n : ta.hi[1] - 1 - n;
            sy += sfd[rw - 1][n] * (C)rb[q][tj - jb + H];
          }
          dyv[q] = left ? sy : -sy;
        }
      }
      C acc[5];
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] = (C)odf[q] * kA;
      acc[0] += kcv[0] * dxv[2] + kcv[1] * dyv[2];  // vx: d/dx sxx + d/dy sxy
      acc[1] += kcv[0] * dxv[3] + kcv[1] * dyv[3];  // vy: d/dx sxy + d/dy syy
      acc[2] += kcg2[0] * dxv[0] + kcl[1] * dyv[1];
      acc[3] += kcg[0] * dxv[1] + kcg[1] * dyv[0];
      acc[4] += kcl[0] * dxv[0] + kcg2[1] * dyv[1];
      C fo[5];
#pragma unroll
      for (int q = 0; q < 4; q++) fo[q] = (C)win[q][r];
      fo[4] = (C)own;
      const int c2[3] = {i, j, 0};
      // external SAT: rows (v_d, s_dd, s_aa) x cols (v_d, s_dd), rows (v_a, sxy) x cols (v_a, sxy)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if (!ta.face_active[e]) continue;
        const int d = e >> 1, a = 1 - d;
        if (c2[d] != ((e & 1) ? ta.hi[d] - 1 : ta.lo[d])) continue;
        const int sdd = d == 0 ? 2 : 4, saa = d == 0 ? 4 : 2;
        const C* m = reinterpret_cast<const C*>(ta.fcoef) + e * 10;
        const C vd = fo[d], sd = fo[sdd], va = fo[a], sa = fo[3];
        acc[d] += m[0] * vd + m[1] * sd;
        acc[sdd] += m[2] * vd + m[3] * sd;
        acc[saa] += m[4] * vd + m[5] * sd;
        acc[a] += m[6] * va + m[7] * sa;
        acc[3] += m[8] * va + m[9] * sa;
      }
      // interface SAT terms written by the fault kernel (interface.cpp:826-962)
      for (int e = 0; e < ta.nif; e++) {
        if (c2[ta.if_dir[e]] != ta.if_fixed[e]) continue;
        const long long o = c2[ta.if_a[e]] - ta.if_orga[e];
        const long long n = ta.if_n[e];
        const double* sn = ta.if_sat[e] + (long long)(ta.if_side[e] * 2 + 0) * 5 * n;
        const double* sv = sn + 5 * n;
#pragma unroll
        for (int q = 0; q < 5; q++) acc[q] -= (C)(sn[q * n + o] + sv[q * n + o]);
      }
      constexpr int s0i[5] = {-1, -1, 0, 1, 2};
#pragma unroll
      for (int q = 0; q < 5; q++) {
        rds.st(q, (R)acc[q], colb, plb);
        C v = fo[q] + kB * acc[q];
        if (q >= 2 && ta.roundtrip) {  // set_stress / remove_stress rounding (FP64 only)
          const long long id = (long long)i * ta.Ny + j;
          double w = (double)v;
          w += ta.s0[s0i[q]];
          if (ta.hetstress) w += ta.hs[(q - 2) * N + id];
          w -= ta.s0[s0i[q]];
          if (ta.hetstress) w -= ta.hs[(q - 2) * N + id];
          v = (C)w;
        }
        rn.st(q, (R)v, colb, plb);
      }
    }
#pragma unroll
    for (int n = 0; n < W - 1; n++)
#pragma unroll
      for (int q = 0; q < 4; q++) win[q][n] = win[q][n + 1];
  }
  lds_barrier();  // a persistent workgroup's next tile reuses the buffers
}

// one block: workgroup b is tile b in XCD-aware order
template <typename R, typename C, int S, int TW, bool WIDE = false>
__global__ void __launch_bounds__(TW) bulk_tiled2d(Tiled2D ta, int par) {
  const int nxcd = 8;
  const int per = (ta.ntiles + nxcd - 1) / nxcd;
  const int tile_id = (blockIdx.x % nxcd) * per + blockIdx.x / nxcd;
  if (tile_id >= ta.ntiles) return;
  tile2d<R, C, S, TW, WIDE>(ta, tile_id, par, ta.dfmode);
}

// ensemble: a persistent grid walks the work list (block descriptor, tile)
// of every instance; uneven instance sizes only change the list length
// (the grid is a multiple of 8: workgroup b runs on XCD b % 8, and the items
// a round of workgroups takes are mapped so that each XCD gets a contiguous
// run of them, neighbouring y tiles sharing that XCD's L2)
template <typename R, typename C, int S, int TW, bool WIDE = false>
__global__ void __launch_bounds__(TW) bulk_tiled2d_batch(const Tiled2D* __restrict__ descs,
                                                        const int2* __restrict__ work, int nwork, int par,
                                                        int dfm) {
  const int first = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
  for (int w = first; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    tile2d<R, C, S, TW, WIDE>(descs[it.x], it.y, par, dfm);
  }
}

// ---------------------------------------------------------------------------
// Curvilinear 2-D mode II block (elastic, or plastic with szz), homogeneous
// material, no dissipation.  Same tiling as tile2d (TW y points per
// workgroup, x planes marched with a register window, y taps from an LDS
// row), but metric and Jacobian vary per point.  block.cpp:1048-1278 writes
// the velocity rates in conservative form, (dt/(rho dx_d J_i)) sum_n c_n J_n
// (m_d . s)_n, and the stress rates with the point's own metric row times
// the velocity derivative.  So the x window carries vx, vy and the x fluxes
// Fx = J (m00 sxx + m01 sxy), Fy = J (m00 sxy + m01 syy) of each plane, and
// the LDS row carries vx, vy and the y fluxes Gx = J (m10 sxx + m11 sxy),
// Gy = J (m10 sxy + m11 syy).  With PL the szz rate is nu (rate sxx + rate
// syy), assigned before the SAT terms (block.cpp:1280-1295).  Points on
// external faces add the SAT of boundary.cpp:265-492 through its linear map
// at that point (ta.smap, built by face_map_kernel from ext_sat).
template <typename R, typename C, int S, int TW, bool PL, bool WIDE>
__device__ __forceinline__ void curv2d_tile(const Tiled2D& ta, const BlockDev& bd, const R* __restrict__ mt,
                                            const R* __restrict__ jac, int tile_id, int par, int dfm) {
#pragma clang fp contract(fast)
  constexpr int NF = PL ? 6 : 5;
  constexpr int r = S - 1, W = 2 * r + 1;
  constexpr int NRB = 2 * (S - 1), NTB = 3 * (S - 1);
  constexpr int H = NTB;            // LDS halo per side
  constexpr int LW = TW + 2 * H;
  __shared__ C row[2][4][LW];       // vx vy Gx Gy, double-buffered by plane parity
  __shared__ C sfd[NRB][NTB];
  const R* __restrict__ f = reinterpret_cast<const R*>(ta.f[par]);
  R* __restrict__ fn = const_cast<R*>(reinterpret_cast<const R*>(ta.f[par ^ 1]));
  R* __restrict__ df = reinterpret_cast<R*>(ta.df);
  const int t = threadIdx.x;
  if (t < NRB * NTB) sfd[t / NTB][t % NTB] = (C)c_fd[1 + t / NTB][t % NTB];
  const int yt = tile_id % ta.nyt, xc = tile_id / ta.nyt;
  const int jb = ta.lo[1] + yt * TW, j = jb + t;
  const bool active = j < ta.hi[1];
  const int ib = ta.lo[0] + xc * ta.cx, ie = min(ib + ta.cx, ta.hi[0]);
  const long long N = ta.N;
  const int jc = min(j, ta.hi[1] - 1);
  const unsigned colb = (unsigned)jc * (unsigned)sizeof(R);
  const unsigned psb = (unsigned)ta.Ny * (unsigned)sizeof(R);
  // df/f traffic a stage does not need: descriptors with no records (loads
  // return 0, stores are dropped, no memory request)
  typedef typename FieldsOf<R, WIDE>::type FA;
  const FA rf(f, N, NF, true), rd(df, N, NF, dfm & DF_LOAD), rds(df, N, NF, dfm & DF_STORE),
      rn(fn, N, NF, dfm & F_STORE);
  const FA rm(mt, N, 4, true), rj(jac, N, 1, true);
  const bool yin = j >= ta.mc[1] && j < ta.mrb[1];
  C kv[2], kg2l[2], kg[2], kl[2];
#pragma unroll
  for (int d = 0; d < 2; d++) {
    kv[d] = (C)bd.K[d][0];
    kg2l[d] = (C)bd.K[d][1];
    kg[d] = (C)bd.K[d][2];
    kl[d] = (C)bd.K[d][3];
  }
  const C kA = (C)ta.A, kB = (C)ta.B, knu = (C)bd.nu;
  C cd[r + 1];
#pragma unroll
  for (int n = 0; n <= r; n++) cd[n] = (C)c_fd[0][r + n];
  auto xplane = [&](int i) { return (unsigned)min(max(i, ta.lo[0]), ta.hi[0] - 1) * psb; };
  // x quantities of one plane at this column: vx, vy, Fx, Fy
  auto xq = [&](unsigned pw, C* o) {
    const C vx = (C)rf.ld(0, colb, pw), vy = (C)rf.ld(1, colb, pw);
    const C sxx = (C)rf.ld(2, colb, pw), sxy = (C)rf.ld(3, colb, pw);
    const C syy = (C)rf.ld(4, colb, pw);
    const C J = (C)rj.ld(0, colb, pw), m0 = (C)rm.ld(0, colb, pw), m1 = (C)rm.ld(1, colb, pw);
    o[0] = vx;
    o[1] = vy;
    o[2] = J * (m0 * sxx + m1 * sxy);
    o[3] = J * (m0 * sxy + m1 * syy);
  };
  // halo entry of this thread (threads 0..2H-1): row position and y index
  const bool hal = t < 2 * H;
  const int hpos = t < H ? t : TW + t;
  const int hj = min(max(jb - H + hpos, ta.lo[1]), ta.hi[1] - 1);
  const unsigned hoff = (unsigned)hj * (unsigned)sizeof(R);
  C win[4][W];
#pragma unroll
  for (int n = 0; n < W - 1; n++) {
    C o[4];
    xq(xplane(ib - r + n), o);
#pragma unroll
    for (int q = 0; q < 4; q++) win[q][n] = o[q];
  }
  // loads in flight for the next plane: the window's leading plane (vx vy
  // sxx sxy syy J m00 m01), this column's own values (sxx sxy syy [szz] J
  // m00 m01 m10 m11) and the halo entry (vx vy sxx sxy syy J m10 m11)
  R nlead[8], nown[9], nhal[8];
  auto issue = [&](int ip) {
    const unsigned pw = xplane(ip + r), pp = (unsigned)ip * psb;
#pragma unroll
    for (int q = 0; q < 5; q++) nlead[q] = rf.ld(q, colb, pw);
    nlead[5] = rj.ld(0, colb, pw);
    nlead[6] = rm.ld(0, colb, pw);
    nlead[7] = rm.ld(1, colb, pw);
#pragma unroll
    for (int q = 0; q < 3; q++) nown[q] = rf.ld(2 + q, colb, pp);
    nown[3] = PL ? rf.ld(5, colb, pp) : (R)0;
    nown[4] = rj.ld(0, colb, pp);
#pragma unroll
    for (int q = 0; q < 4; q++) nown[5 + q] = rm.ld(q, colb, pp);
    if (hal) {
#pragma unroll
      for (int q = 0; q < 5; q++) nhal[q] = rf.ld(q, hoff, pp);
      nhal[5] = rj.ld(0, hoff, pp);
      nhal[6] = rm.ld(2, hoff, pp);
      nhal[7] = rm.ld(3, hoff, pp);
    }
  };
  issue(ib);
  for (int i = ib; i < ie; i++) {
    auto* const rb = row[(i - ib) & 1];  // this plane's LDS row buffer
    const unsigned plb = (unsigned)i * psb;
    {
      const C J = (C)nlead[5], m0 = (C)nlead[6], m1 = (C)nlead[7];
      const C sxx = (C)nlead[2], sxy = (C)nlead[3], syy = (C)nlead[4];
      win[0][W - 1] = (C)nlead[0];
      win[1][W - 1] = (C)nlead[1];
      win[2][W - 1] = J * (m0 * sxx + m1 * sxy);
      win[3][W - 1] = J * (m0 * sxy + m1 * syy);
    }
    const C sxx = (C)nown[0], sxy = (C)nown[1], syy = (C)nown[2], szz = (C)nown[3];
    const C Ji = (C)nown[4], m00 = (C)nown[5], m01 = (C)nown[6], m10 = (C)nown[7], m11 = (C)nown[8];
    rb[0][H + t] = win[0][r];
    rb[1][H + t] = win[1][r];
    rb[2][H + t] = Ji * (m10 * sxx + m11 * sxy);
    rb[3][H + t] = Ji * (m10 * sxy + m11 * syy);
    if (hal) {
      const C J = (C)nhal[5], m0 = (C)nhal[6], m1 = (C)nhal[7];
      rb[0][hpos] = (C)nhal[0];
      rb[1][hpos] = (C)nhal[1];
      rb[2][hpos] = J * (m0 * (C)nhal[2] + m1 * (C)nhal[3]);
      rb[3][hpos] = J * (m0 * (C)nhal[3] + m1 * (C)nhal[4]);
    }
    lds_barrier();
    if (i + 1 < ie) issue(i + 1);
    R odf[5];  // (the szz rate is assigned, so its old value is not read)
#pragma unroll
    for (int q = 0; q < 5; q++) odf[q] = rd.ld(q, colb, plb);
    if (active) {
      C dxv[4], dyv[4];  // d/dx of vx vy Fx Fy; d/dy of vx vy Gx Gy
      if (i >= ta.mc[0] && i < ta.mrb[0]) {
#pragma unroll
        for (int q = 0; q < 4; q++) dxv[q] = central<C, r>(cd, [&](int n) { return win[q][n]; });
      } else {  // one-sided row along x: taps are whole planes (wave-uniform)
        const bool left = i - ta.lo[0] < NRB;
        const int rw = left ? i - ta.lo[0] + 1 : ta.hi[0] - i;
#pragma unroll
        for (int n = 0; n < NTB; n++) {
          const int tap = left ? ta.lo[0] + n : ta.hi[0] - 1 - n;
          const C c = (C)c_fd[rw][n];
          C o[4];
          xq((unsigned)tap * psb, o);
#pragma unroll
          for (int q = 0; q < 4; q++) dxv[q] = n == 0 ? c * o[q] : dxv[q] + c * o[q];
        }
        if (!left)
#pragma unroll
          for (int q = 0; q < 4; q++) dxv[q] = -dxv[q];
      }
      if (yin) {
#pragma unroll
        for (int q = 0; q < 4; q++) dyv[q] = central<C, r>(cd, [&](int n) { return rb[q][t + H - r + n]; });
      } else {  // one-sided row along y: taps in the LDS row
        const bool left = j - ta.lo[1] < NRB;
        const int rw = left ? j - ta.lo[1] + 1 : ta.hi[1] - j;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          C sy = 0.;
#pragma unroll
          for (int n = 0; n < NTB; n++) {
            const int tj = left ? ta.lo[1] + n : ta.hi[1] - 1 - n;
            sy += sfd[rw - 1][n] * rb[q][tj - jb + H];
          }
          dyv[q] = left ? sy : -sy;
        }
      }
      C acc[6];
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] = (C)odf[q] * kA;
      acc[0] += (kv[0] * dxv[2] + kv[1] * dyv[2]) / Ji;
      acc[1] += (kv[0] * dxv[3] + kv[1] * dyv[3]) / Ji;
      acc[2] += kg2l[0] * m00 * dxv[0] + kl[0] * m01 * dxv[1] + kg2l[1] * m10 * dyv[0] + kl[1] * m11 * dyv[1];
      acc[3] += kg[0] * (m00 * dxv[1] + m01 * dxv[0]) + kg[1] * (m10 * dyv[1] + m11 * dyv[0]);
      acc[4] += kg2l[0] * m01 * dxv[1] + kl[0] * m00 * dxv[0] + kg2l[1] * m11 * dyv[1] + kl[1] * m10 * dyv[0];
      acc[5] = PL ? knu * (acc[2] + acc[4]) : (C)0;
      const long long id = (long long)i * ta.Ny + j;
      const int c2[2] = {i, j};
      const C fo[6] = {win[0][r], win[1][r], sxx, sxy, syy, szz};
      // external boundaries of this block, in face order (block.cpp:429-440)
#pragma unroll
      for (int fc = 0; fc < 4; fc++) {
        if (!bd.face[fc].active) continue;
        const int d = fc >> 1, a = 1 - d;
        if (c2[d] != ((fc & 1) ? bd.hi[d] - 1 : bd.lo[d])) continue;
        const R* m = reinterpret_cast<const R*>(ta.smap[fc]) + (long long)(c2[a] - bd.lo[a]) * (NF * 5);
#pragma unroll
        for (int p = 0; p < NF; p++) {
          C sp = (C)m[p * 5] * fo[0];
#pragma unroll
          for (int q = 1; q < 5; q++) sp += (C)m[p * 5 + q] * fo[q];
          acc[p] += sp;
        }
      }
      // interface SAT terms written by the fault kernel (interface.cpp:826-962)
      for (int e = 0; e < ta.nif; e++) {
        const int cd_ = ta.if_dir[e] == 0 ? i : j, ca = ta.if_a[e] == 0 ? i : j;
        if (cd_ != ta.if_fixed[e]) continue;
        const long long o = ca - ta.if_orga[e];
        const long long n = ta.if_n[e];
        const double* sn = ta.if_sat[e] + (long long)(ta.if_side[e] * 2 + 0) * NF * n;
        const double* sv = sn + NF * n;
#pragma unroll
        for (int q = 0; q < NF; q++) acc[q] -= (C)(sn[q * n + o] + sv[q * n + o]);
      }
      constexpr int s0i[5] = {-1, -1, 0, 1, 2};
#pragma unroll
      for (int q = 0; q < NF; q++) {
        rds.st(q, (R)acc[q], colb, plb);
        C v = fo[q] + kB * acc[q];
        if (!PL && q >= 2 && ta.roundtrip) {  // set_stress / remove_stress rounding (FP64 elastic only)
          double w = (double)v;
          w += ta.s0[s0i[q]];
          if (ta.hetstress) w += ta.hs[(q - 2) * N + id];
          w -= ta.s0[s0i[q]];
          if (ta.hetstress) w -= ta.hs[(q - 2) * N + id];
          v = (C)w;
        }
        rn.st(q, (R)v, colb, plb);
      }
    }
#pragma unroll
    for (int n = 0; n < W - 1; n++)
#pragma unroll
      for (int q = 0; q < 4; q++) win[q][n] = win[q][n + 1];
  }
  lds_barrier();  // a persistent workgroup's next tile reuses the buffers
}

// one block: workgroup b is tile b in XCD-aware order
template <typename R, typename C, int S, int TW, bool PL, bool WIDE = false>
__global__ void __launch_bounds__(TW) bulk_curv2d(Tiled2D ta, BlockDev bd, const R* __restrict__ mt,
                                                  const R* __restrict__ jac, int par) {
  const int nxcd = 8;
  const int per = (ta.ntiles + nxcd - 1) / nxcd;
  const int tile_id = (blockIdx.x % nxcd) * per + blockIdx.x / nxcd;
  if (tile_id >= ta.ntiles) return;
  curv2d_tile<R, C, S, TW, PL, WIDE>(ta, bd, mt, jac, tile_id, par, ta.dfmode);
}

// every block of a problem in one launch: work item (block, tile), one
// workgroup per item (grid a multiple of 8, XCD-grouped as in
// bulk_tiled2d_batch)
template <typename R, typename C, int S, int TW, bool PL, bool WIDE = false>
__global__ void __launch_bounds__(TW) bulk_curv2d_multi(const Tiled2D* __restrict__ descs,
                                                        const BlockDev* __restrict__ bds, const int2* __restrict__ work,
                                                        int nwork, const R* __restrict__ mt, const R* __restrict__ jac,
                                                        int par, int dfm) {
  const int w = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
  if (w >= nwork) return;
  const int2 it = work[w];
  curv2d_tile<R, C, S, TW, PL, WIDE>(descs[it.x], bds[it.x], mt, jac, it.y, par, dfm);
}

// ensemble fault update: work item = (instance, first fault point of a
// chunk of 256); the f arrays of the stage parity come with the instance
struct FaultBatchDev {
  FaultDev F;
  GridDev g;
  const void* f[2];
  const void* mat;
  const double* hs;
  double dt;  // this instance's time step (stage times come from the step counter)
};

// LAW: the friction law every instance shares (-1: any, chosen per point);
// slip weakening takes fault_point's lean form (fewer registers, so more
// waves in flight over the ensemble's faults)
template <int M, typename R, int LAW = -1>
__global__ void __launch_bounds__(256) fault_batch_kernel(const FaultBatchDev* __restrict__ inst,
                                                          const int2* __restrict__ work, int nwork,
                                                          FaultStageDev fs, int par, int* __restrict__ err) {
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    const FaultBatchDev& I = inst[it.x];
    const int p = it.y + threadIdx.x;
    FaultStageDev fsi = fs;
    fsi.dt = I.dt;
    if (p < I.F.n)
      fault_point<M, R, LAW, NoHook, LAW == 2>(I.F, I.g, fsi, reinterpret_cast<const R*>(I.f[par]),
                                               reinterpret_cast<const R*>(I.mat), I.hs, err, nullptr, p, nullptr);
  }
}

// A 2-D RK stage in one launch, fault solve included.  A y-interface between
// two rectilinear blocks whose x-chunks line up is cut into the x-chunks of
// the tiles: work item (-1 - k, 0) is chunk k, one workgroup that solves the
// chunk's fault points (one lane each, fault_point as in fault_kernel), then
// updates the two tiles that hold the interface rows, reading the SAT terms
// it has just written.  Work items (desc, tile) are the other tiles.  No
// other workgroup reads those SAT terms or this chunk's fault state, so the
// stage needs no separate fault launch and no grid-wide ordering; every
// fault point and bulk point is computed once, with the arithmetic of the
// separate launches (domain.cpp:228-305: the interface solve before the
// bulk update).
struct FaultPair2D {
  int fd;            // fault descriptor (FaultBatchDev) of the interface
  int d0, d1;        // block descriptors of side 1 and side 2
  int tile0, tile1;  // tile ids (yt + xc nyt) holding each side's interface row
  int p0, np;        // first fault point of the chunk and its point count
};

// The fault part (fault_point's lean form) sets the kernel's register count,
// about 135 VGPRs (3 waves per SIMD).  Capping it lower spills to scratch,
// and the kernels with scratch measured slower than the uncapped kernel.
template <typename R, typename C, int S, int TW, int LAW>
__global__ void __launch_bounds__(TW) bulk_fault2d_batch(const Tiled2D* __restrict__ descs, const int2* __restrict__ work, int nwork, int par, int dfm,
                   const FaultBatchDev* __restrict__ fdv, const FaultPair2D* __restrict__ pairs, FaultStageDev fs,
                   int* __restrict__ err) {
  const int first = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
  for (int w = first; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    int nt = 1, d0 = it.x, t0 = it.y, d1 = 0, t1 = 0;
    if (it.x < 0) {
      const FaultPair2D pr = pairs[-1 - it.x];
      if ((int)threadIdx.x < pr.np) {
        const FaultBatchDev& I = fdv[pr.fd];
        FaultStageDev fsi = fs;
        fsi.dt = I.dt;
        fault_point<0, R, LAW, NoHook, true>(I.F, I.g, fsi, reinterpret_cast<const R*>(I.f[par]),
                                             reinterpret_cast<const R*>(I.mat), I.hs, err, nullptr,
                                             pr.p0 + (int)threadIdx.x, nullptr);
      }
      __syncthreads();  // the SAT terms in global memory, visible to the workgroup
      nt = 2;
      d0 = pr.d0; t0 = pr.tile0;
      d1 = pr.d1; t1 = pr.tile1;
    }
    for (int k = 0; k < nt; k++) tile2d<R, C, S, TW, false>(descs[k ? d1 : d0], k ? t1 : t0, par, dfm);
  }
}

// Linear map of the external SAT on one face point: column q is ext_sat
// applied to the unit state e_q (9 fields), so M . f reproduces the penalty
// added to df for state f at that point.  Run once at set-up.
template <int M>
__global__ void face_matrix_kernel(FaceDev F, GridDev g, double dt, int u, int v, const double* __restrict__ unit,
                                   double* __restrict__ out) {
  constexpr int NF = Nf<M>::v;
  const int q = threadIdx.x;
  if (q >= NF) return;
  GridDev gg = g;
  gg.N = NF;
  gg.hetmat = 0;
  double acc[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  ext_sat<M, double, double>(acc, F, unit, nullptr, gg, q, u, v, dt);
  for (int p = 0; p < NF; p++) out[p * NF + q] = acc[p];
}

// Linear maps of the external SAT at every point of a face: out[u][p][q] is
// the rate of field p at face point u per unit of field q (q over the mode's
// fields; p over nr rows, which include szz for plastic mode II).
template <int M>
__global__ void face_map_kernel(FaceDev F, GridDev g, double dt, int n0, int nr, const double* __restrict__ unit,
                                double* __restrict__ out) {
  constexpr int NC = Nf<M>::v;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n0 * NC) return;
  const int u = t / NC, q = t % NC;
  GridDev gg = g;
  gg.N = NC;
  gg.hetmat = 0;
  double acc[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  ext_sat<M, double, double>(acc, F, unit, nullptr, gg, q, u, 0, dt);
  for (int p = 0; p < nr; p++) out[((long long)u * nr + p) * NC + q] = acc[p];
}

// ---------------------------------------------------------------------------
// Device-side grid set-up (block.cpp:769-1046, block.cpp:368-403): the
// transfinite interpolation of a block's face surfaces, the derivatives
// dX_c/dxi_d by the SBP operator along grid lines (one-sided rows near block
// edges), the Jacobian and the inverse metric.  The expressions and their
// order are those of the host set-up (FP64, no contraction), so the arrays
// are bit-identical to it.
struct SurfDev {  // a face surface [ndim][n0][n1]
  const double* p;
  int n0, n1;
};
struct GridBlockDev {
  SurfDev F[6];
  int nx[3], xm[3];
  double dx[3];
  int ndim, Ny, Nz;
  long long N;
};

__device__ __forceinline__ double surf_at(const SurfDev& s, int c, int i, int j) {
  return s.p[((long long)c * s.n0 + i) * s.n1 + j];
}

__global__ void __launch_bounds__(256) grid_tfi_kernel(GridBlockDev g, double* __restrict__ x) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nb = (long long)g.nx[0] * g.nx[1] * g.nx[2];
  if (t >= nb) return;
  const int k = (int)(t % g.nx[2]);
  const int j = (int)((t / g.nx[2]) % g.nx[1]);
  const int i = (int)(t / ((long long)g.nx[1] * g.nx[2]));
  const int ny = g.nx[1], nz = g.nx[2];
  const SurfDev* F = g.F;
  const long long gi = ((long long)(g.xm[0] + i) * g.Ny + (g.xm[1] + j)) * g.Nz + (g.xm[2] + k);
  for (int c = 0; c < g.ndim; c++) {
    const double p = (double)i * g.dx[0], q = (double)j * g.dx[1], r = (double)k * g.dx[2];
    double v = ((1. - p) * surf_at(F[0], c, j, k) + p * surf_at(F[1], c, j, k) + (1. - q) * surf_at(F[2], c, i, k) +
                q * surf_at(F[3], c, i, k));
    if (g.ndim == 3) v += (1. - r) * surf_at(F[4], c, i, j) + r * surf_at(F[5], c, i, j);
    v -= ((1. - q) * (1. - p) * surf_at(F[0], c, 0, k) + (1. - q) * p * surf_at(F[1], c, 0, k) +
          q * (1. - p) * surf_at(F[0], c, ny - 1, k) + q * p * surf_at(F[1], c, ny - 1, k));
    if (g.ndim == 3) {
      v -= ((1. - p) * (1. - r) * surf_at(F[0], c, j, 0) + p * (1. - r) * surf_at(F[1], c, j, 0) +
            (1. - q) * (1. - r) * surf_at(F[2], c, i, 0) + q * (1. - r) * surf_at(F[3], c, i, 0) +
            (1. - p) * r * surf_at(F[0], c, j, nz - 1) + p * r * surf_at(F[1], c, j, nz - 1) +
            (1. - q) * r * surf_at(F[2], c, i, nz - 1) + q * r * surf_at(F[3], c, i, nz - 1));
      v += ((1. - p) * (1. - q) * (1. - r) * surf_at(F[0], c, 0, 0) + p * (1. - q) * (1. - r) * surf_at(F[1], c, 0, 0) +
            (1. - p) * q * (1. - r) * surf_at(F[0], c, ny - 1, 0) + (1. - p) * (1. - q) * r * surf_at(F[0], c, 0, nz - 1) +
            p * q * (1. - r) * surf_at(F[1], c, ny - 1, 0) + p * (1. - q) * r * surf_at(F[1], c, 0, nz - 1) +
            (1. - p) * q * r * surf_at(F[0], c, ny - 1, nz - 1) + p * q * r * surf_at(F[1], c, ny - 1, nz - 1));
    }
    x[(long long)c * g.N + gi] = v;
  }
}

// metric derivatives, Jacobian and inverse metric of one block (SBP table
// from c_fd, sbporder s); *bad is set where the Jacobian is not positive
__global__ void __launch_bounds__(256) grid_metric_kernel(GridBlockDev g, int s, const double* __restrict__ x,
                                                          double* __restrict__ metric, double* __restrict__ jac,
                                                          int* __restrict__ bad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nb = (long long)g.nx[0] * g.nx[1] * g.nx[2];
  if (t >= nb) return;
  const int loc[3] = {(int)(t / ((long long)g.nx[1] * g.nx[2])), (int)((t / g.nx[2]) % g.nx[1]), (int)(t % g.nx[2])};
  const long long gstride[3] = {(long long)g.Ny * g.Nz, (long long)g.Nz, 1};
  const long long gi = (g.xm[0] + loc[0]) * gstride[0] + (long long)(g.xm[1] + loc[1]) * gstride[1] + (g.xm[2] + loc[2]);
  double J9[3][3];
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++) J9[c][d] = (c == d) ? 1. : 0.;
  const int edge = 2 * (s - 1);
  for (int d = 0; d < g.ndim; d++) {
    const int pos = loc[d], n = g.nx[d];
    const long long line0 = gi - (long long)pos * gstride[d];
    for (int c = 0; c < g.ndim; c++) {
      const double* X = x + (long long)c * g.N + line0;
      double acc = 0.;
      if (pos < edge) {
        for (int tp = 0; tp < 3 * (s - 1); tp++) acc += c_fd[pos + 1][tp] * X[(long long)tp * gstride[d]] / g.dx[d];
      } else if (pos < n - edge) {
        for (int tp = 0; tp < 2 * s - 1; tp++)
          acc += c_fd[0][tp] * X[(long long)(pos - s + 1 + tp) * gstride[d]] / g.dx[d];
      } else {
        for (int tp = 0; tp < 3 * (s - 1); tp++) acc -= c_fd[n - pos][tp] * X[(long long)(n - 1 - tp) * gstride[d]] / g.dx[d];
      }
      J9[c][d] = acc;
    }
  }
  auto X = [&](int c, int d) { return J9[c][d]; };
  const double jc = (X(0, 0) * (X(1, 1) * X(2, 2) - X(1, 2) * X(2, 1)) - X(1, 0) * (X(0, 1) * X(2, 2) - X(0, 2) * X(2, 1)) +
                     X(2, 0) * (X(0, 1) * X(1, 2) - X(0, 2) * X(1, 1)));
  if (!(jc > 0.)) atomicOr(bad, 1);
  jac[gi] = jc;
  for (int l = 0; l < g.ndim; l++)
    for (int m = 0; m < g.ndim; m++)
      metric[(long long)(l * g.ndim + m) * g.N + gi] =
          ((X((m + 1) % 3, (l + 1) % 3) * X((m + 2) % 3, (l + 2) % 3) - X((m + 1) % 3, (l + 2) % 3) * X((m + 2) % 3, (l + 1) % 3)) / jc);
}

// grid values at n global points: x[q][ndim], metric[q][ndim^2], jac[q]
__global__ void grid_gather_kernel(const long long* __restrict__ idx, int n, int ndim, long long N,
                                   const double* __restrict__ x, const double* __restrict__ metric,
                                   const double* __restrict__ jac, double* __restrict__ ox, double* __restrict__ om,
                                   double* __restrict__ oj) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long long g = idx[q];
  for (int c = 0; c < ndim; c++) ox[(long long)q * ndim + c] = x[c * N + g];
  for (int lm = 0; lm < ndim * ndim; lm++) om[(long long)q * ndim * ndim + lm] = metric[lm * N + g];
  oj[q] = jac[g];
}

// max over the block box [lo, hi) and directions l of cs |grad xi_l| / dx_l
// (block.cpp:368-403): the values are non-negative, so the largest bit
// pattern is the largest value
__global__ void __launch_bounds__(256) grid_worst_kernel(int lo0, int lo1, int lo2, int n0, int n1, int n2, int ndim,
                                                         int Ny, int Nz, long long N, double dx0, double dx1, double dx2,
                                                         double cs, const double* __restrict__ metric,
                                                         const double* __restrict__ mat,
                                                         unsigned long long* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nb = (long long)n0 * n1 * n2;
  double w = 0.;
  if (t < nb) {
    const int k = (int)(t % n2), j = (int)((t / n2) % n1), i = (int)(t / ((long long)n1 * n2));
    const long long g = ((long long)(lo0 + i) * Ny + (lo1 + j)) * Nz + (lo2 + k);
    const double dx[3] = {dx0, dx1, dx2};
    for (int l = 0; l < ndim; l++) {
      double v = 0.;
      for (int m = 0; m < ndim; m++) {
        const double q = metric[(long long)(l * ndim + m) * N + g];
        v += q * q;
      }
      const double c = mat ? sqrt(mat[2 * N + g] / mat[g]) : cs;
      v = c * sqrt(v) / dx[l];
      if (v > w) w = v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) w = fmax(w, __shfl_xor(w, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(w));
}

// whether a block's metric is diagonal and uniform and its Jacobian uniform
// (the host check of block_metric_uniform: 1e-9 relative of the centre
// point's values); *bad is set otherwise
__global__ void __launch_bounds__(256) grid_uniform_kernel(int lo0, int lo1, int lo2, int n0, int n1, int n2, int ndim,
                                                           int Ny, int Nz, long long N, double jc, double md0,
                                                           double md1, double md2, const double* __restrict__ metric,
                                                           const double* __restrict__ jac, int* __restrict__ bad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nb = (long long)n0 * n1 * n2;
  if (t >= nb) return;
  const int k = (int)(t % n2), j = (int)((t / n2) % n1), i = (int)(t / ((long long)n1 * n2));
  const long long g = ((long long)(lo0 + i) * Ny + (lo1 + j)) * Nz + (lo2 + k);
  const double mdiag[3] = {md0, md1, md2};
  bool b = fabs(jac[g] - jc) > 1e-9 * fabs(jc);
  for (int l = 0; l < ndim; l++)
    for (int m = 0; m < ndim; m++) {
      const double v = metric[(long long)(l * ndim + m) * N + g];
      if (l == m ? fabs(v - mdiag[l]) > 1e-9 * fabs(mdiag[l]) : fabs(v) > 1e-9 * fabs(mdiag[l])) b = true;
    }
  if (b) atomicOr(bad, 1);
}

}  // namespace fdf
