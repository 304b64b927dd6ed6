// Host-side problem compiler (see setup.hpp).  Single address space, one GPU:
// the whole grid is one global SoA box, blocks are sub-boxes of it, exactly
// the reference's single-rank layout (fields.cpp:87-98, block.cpp:161-179).
#include "setup.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace fdf {

namespace {

[[noreturn]] void fail(const std::string& m) { throw Error{m}; }

// Token scanner with the semantics of the reference's getline/operator>>
// section scans (e.g. domain.cpp:33-71): a section starts after a line equal
// to its header; missing headers are errors unless optional.
class Scanner {
 public:
  explicit Scanner(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) fail("Error opening input file " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    buf_ = ss.str();
  }
  bool seek(const std::string& header, bool from_start = true) {
    if (from_start) pos_ = 0;
    while (pos_ < buf_.size()) {
      size_t e = buf_.find('\n', pos_);
      bool last = (e == std::string::npos);
      std::string line = buf_.substr(pos_, last ? std::string::npos : e - pos_);
      pos_ = last ? buf_.size() : e + 1;
      if (line == header) return !last;
    }
    return false;
  }
  bool line(std::string& out) {
    if (pos_ >= buf_.size()) return false;
    size_t e = buf_.find('\n', pos_);
    if (e == std::string::npos) {
      out = buf_.substr(pos_);
      pos_ = buf_.size();
    } else {
      out = buf_.substr(pos_, e - pos_);
      pos_ = e + 1;
    }
    return true;
  }
  std::string word() {
    while (pos_ < buf_.size() && isspace((unsigned char)buf_[pos_])) pos_++;
    if (pos_ >= buf_.size()) fail("unexpected end of input file");
    size_t b = pos_;
    while (pos_ < buf_.size() && !isspace((unsigned char)buf_[pos_])) pos_++;
    return buf_.substr(b, pos_ - b);
  }
  double real() {
    std::string w = word();
    char* e = nullptr;
    double v = strtod(w.c_str(), &e);
    if (e == w.c_str()) fail("bad number '" + w + "' in input file");
    return v;
  }
  int integer() {
    std::string w = word();
    char* e = nullptr;
    long v = strtol(w.c_str(), &e, 10);
    if (e == w.c_str()) fail("bad integer '" + w + "' in input file");
    return (int)v;
  }

 private:
  std::string buf_;
  size_t pos_ = 0;
};

// a binary input file of n doubles; open_msg: the reference's message when
// the file cannot be opened (fields.cpp:497,627, friction.cpp:555,627,
// slipweak.cpp:358, stz.cpp:406, surface.cpp:44)
std::vector<double> load_f8(const std::string& path, size_t n, const std::string& open_msg, const char* what) {
  std::vector<double> v(n);
  std::ifstream in(path, std::ios::binary);
  if (!in) fail(open_msg);
  if (!in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double))))
    fail(std::string("Error reading ") + what + " from file " + path);
  return v;
}

void init_sbp(SBP& T, int s) {  // coefficient tables of fd.cpp:52-121
  if (s < 2 || s > 4) fail("Error in initializing FD method");
  T = SBP();
  T.s = s;
  auto put = [](double* r, std::initializer_list<double> v) {
    for (double x : v) *r++ = x;
  };
  if (s == 2) {
    put(T.c[0], {-0.5, 0., 0.5});
    put(T.c[1], {-1., 1., 0.});
    put(T.c[2], {-0.5, 0., 0.5});
    put(T.d[0], {0.5, -1., 0.5});
    put(T.d[1], {-1., 1., 0.});
    put(T.d[2], {0.5, -1., 0.5});
    T.h0 = 0.5;
  } else if (s == 3) {
    put(T.c[0], {1. / 12., -2. / 3., 0., 2. / 3., -1. / 12.});
    put(T.c[1], {-24. / 17., 59. / 34., -4. / 17., -3. / 34., 0., 0.});
    put(T.c[2], {-0.5, 0., 0.5, 0., 0., 0.});
    put(T.c[3], {4. / 43., -59. / 86., 0., 59. / 86., -4. / 43., 0.});
    put(T.c[4], {3. / 98., 0., -59. / 98., 0., 32. / 49., -4. / 49.});
    put(T.d[0], {-1. / 12., 1. / 3., -1. / 2., 1. / 3., -1. / 12.});
    put(T.d[1], {-8. / 17., 16. / 17., -8. / 17., 0., 0., 0.});
    put(T.d[2], {16. / 59., -36. / 59., 24. / 59., -4. / 59., 0., 0.});
    put(T.d[3], {-8. / 43., 24. / 43., -28. / 43., 16. / 43., -4. / 43, 0.});
    put(T.d[4], {0., -4. / 49., 16. / 49., -24. / 49., 16. / 49., -4. / 49.});
    T.h0 = 17. / 48.;
  } else {
    put(T.c[0], {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.});
    put(T.c[1], {-1.582533518939116, 2.033378678700676, -0.141512858744873, -0.450398306578272, 0.104488069284042,
                 0.036577936277544});
    put(T.c[2], {-0.462059195631158, 0., 0.287258622978251, 0.258816087376832, -0.069112065532624,
                 -0.014903449191300});
    put(T.c[3], {0.071247104721830, -0.636451095137907, 0., 0.606235523609147, -0.022902190275815,
                 -0.018129342917256});
    put(T.c[4], {0.114713313798970, -0.290087484386815, -0.306681191361148, 0., 0.520262285050482,
                 -0.051642265516119, 0.013435342414630});
    put(T.c[5], {-0.036210680656541, 0.105400944933782, 0.015764336127392, -0.707905442575989, 0.,
                 0.769199413962647, -0.164529643265203, 0.018281071473911});
    put(T.c[6], {-0.011398193015050, 0.020437334208704, 0.011220896474665, 0.063183694641876, -0.691649024426814, 0.,
                 0.739709139060752, -0.147941827812150, 0.016437980868017});
    put(T.d[0], {1. / 60., -1. / 10., 1. / 4., -1. / 3., 1. / 4., -1. / 10., 1. / 60.});
    put(T.d[1], {-0.105502234595941, 0.316506703787823, -0.316506703787823, 0.105502234595941});
    put(T.d[2], {0.071922084408557, -0.227753267293765, 0.251727295429951, -0.107883126612836, 0.011987014068093});
    put(T.d[3], {-0.159350793065290, 0.557727775728513, -0.743637034304685, 0.478052379195869, -0.159350793065290,
                 0.026558465510882});
    put(T.d[4], {0.026870684829259, -0.120918081731666, 0.241836163463333, -0.282142190707221, 0.201530136219444,
                 -0.080612054487778, 0.013435342414630});
    put(T.d[5], {0., 0.018281071473911, -0.109686428843468, 0.274216072108671, -0.365621429478228, 0.274216072108671,
                 -0.109686428843468, 0.018281071473911});
    put(T.d[6], {0., 0., 0.016437980868017, -0.098627885208100, 0.246569713020251, -0.328759617360334,
                 0.246569713020251, -0.098627885208100, 0.016437980868017});
    T.h0 = 13649. / 43200.;
  }
}

void init_rk(LowStorageRK& R, int order) {  // rk.cpp:17-91
  R = LowStorageRK();
  if (order == 1) {
    R.nstages = 1;
    R.B[0] = 1.;
    R.C[1] = 1.;
  } else if (order == 2) {
    R.nstages = 2;
    R.A[1] = -1.;
    R.B[0] = 1.;
    R.B[1] = 0.5;
    R.C[1] = 1.;
    R.C[2] = 1.;
  } else if (order == 3) {
    R.nstages = 3;
    R.A[1] = -5. / 9.;
    R.A[2] = -153. / 128.;
    R.B[0] = 1. / 3.;
    R.B[1] = 15. / 16.;
    R.B[2] = 8. / 15.;
    R.C[1] = 1. / 3.;
    R.C[2] = 3. / 4.;
    R.C[3] = 1.;
  } else if (order == 4) {  // Carpenter & Kennedy (1994) 2N-storage RK4(5)
    R.nstages = 5;
    R.A[1] = -567301805773. / 1357537059087.;
    R.A[2] = -2404267990393. / 2016746695238.;
    R.A[3] = -3550918686646. / 2091501179385.;
    R.A[4] = -1275806237668. / 842570457699.;
    R.B[0] = 1432997174477. / 9575080441755.;
    R.B[1] = 5161836677717. / 13612068292357.;
    R.B[2] = 1720146321549. / 2090206949498.;
    R.B[3] = 3134564353537. / 4481467310338.;
    R.B[4] = 2277821191437. / 14882151754819.;
    R.C[1] = 1432997174477. / 9575080441755.;
    R.C[2] = 2526269341429. / 6820363962896.;
    R.C[3] = 2006345519317. / 3224310063776.;
    R.C[4] = 2802321613138. / 2924317926251.;
    R.C[5] = 1.;
  } else {
    fail("Error in initializing RK method");
  }
}

// Shape value of one perturbation at fault index (i,j): pert.cpp:38-41
// (coordinates linear in index from the block box) and pert.cpp:59-136.
struct Shape {
  int kind;
  double x0, dx, y0, dy, a, b, c, d;
  double X(int i) const { return b * (double)i + a; }
  double Y(int j) const { return d * (double)j + c; }
  double at(int i, int j) const {
    const double eps = 1.e-14;
    switch (kind) {
      case 0:
        return 1.;
      case 1: {
        double xv = 0., yv = 0.;
        if (!(dx < eps)) { double r = (X(i) - x0) / dx; xv = r * r; }
        if (!(dy < eps)) { double r = (Y(j) - y0) / dy; yv = r * r; }
        return exp(-xv - yv);
      }
      case 2: {
        double xv = 0., yv = 0.;
        if (!(dx < eps)) { double r = (X(i) - x0) / dx; xv = r * r; }
        if (!(dy < eps)) { double r = (Y(j) - y0) / dy; yv = r * r; }
        return (xv + yv <= 1.) ? 1. : 0.;
      }
      case 3: {
        double xv = (dx < eps) ? 1. : X(i) / dx + x0;
        double yv = (dy < eps) ? 1. : Y(j) / dy + y0;
        return xv * yv;
      }
      default: {
        auto box = [eps](double u, double u0, double w) {
          if (w < eps) return 1.;
          if (fabs(fabs(u - u0) - w) < eps) return 0.5;
          if (fabs(u - u0) < w) return 1.;
          return 0.;
        };
        return box(X(i), x0, dx) * box(Y(j), y0, dy);
      }
    }
  }
};

PertH make_pert(Scanner& sc, int namp, const IfaceH& I, const double x2[2], const double l2[2]) {
  std::string kind = sc.word();
  Shape sh;
  if (kind == "constant") sh.kind = 0;
  else if (kind == "gaussian") sh.kind = 1;
  else if (kind == "ellipse") sh.kind = 2;
  else if (kind == "linear") sh.kind = 3;
  else if (kind == "boxcar") sh.kind = 4;
  else fail("unknown perturbation type " + kind);
  PertH p;
  p.t0 = sc.real();
  sh.x0 = sc.real();
  sh.dx = sc.real();
  sh.y0 = sc.real();
  sh.dy = sc.real();
  if (p.t0 < 0. || sh.dx < 0. || sh.dy < 0.) fail("invalid perturbation parameters");
  for (int k = 0; k < namp; k++) p.amp[k] = sc.real();
  const int n[2] = {I.n0, I.n1};
  // single process: local start equals global start, so the offset term is 0/(n-1)
  sh.a = 0. / (double)(n[0] - 1) * l2[0] + x2[0];
  sh.b = l2[0] / (double)(n[0] - 1);
  sh.c = 0. / (double)(n[1] - 1) * l2[1] + x2[1];
  sh.d = l2[1] / (double)(n[1] - 1);
  p.xy.resize((size_t)I.n0 * I.n1);
  for (int i = 0; i < I.n0; i++)
    for (int j = 0; j < I.n1; j++) p.xy[(size_t)i * I.n1 + j] = sh.at(i, j);
  return p;
}

// flat or file surface of a block face (surface.cpp:11-101): [ndim][n0][n1]
std::vector<double> face_surface(const SetupH& S, const BlockH& B, int face, int& n0, int& n1) {
  const int d = face / 2;
  const int a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
  n0 = B.nx[a];
  n1 = B.nx[b];
  const size_t n = (size_t)n0 * n1;
  if (B.bfile[face] != "none")
    return load_f8(B.bfile[face], (size_t)S.ndim * n, "Error reading surface from file " + B.bfile[face], "surface");
  double corner[3] = {B.x0[0], B.x0[1], B.x0[2]};
  if (face & 1) corner[d] = B.x0[d] + B.L[d];
  const double la = B.L[a], lb = B.L[b];
  if (la < 0. || lb < 0.) fail("negative block length");
  std::vector<double> out((size_t)S.ndim * n, 0.);
  for (size_t q = 0; q < n; q++) out[(size_t)d * n + q] = corner[d];
  for (int j = 0; j < n0; j++)
    for (int k = 0; k < n1; k++) {
      out[(size_t)a * n + (size_t)j * n1 + k] = corner[a] + la * (double)(j) / (double)(n0 - 1);
      if (S.ndim == 3) out[(size_t)b * n + (size_t)j * n1 + k] = corner[b] + lb * (double)(k) / (double)(n1 - 1);
    }
  return out;
}

struct SurfView {
  const double* p;
  int n0, n1;
  double operator()(int c, int i, int j) const { return p[((size_t)c * n0 + i) * n1 + j]; }
};

// transfinite interpolation (block.cpp:795-829) and metric by the SBP
// operator (block.cpp:862-1011)
void build_block_grid(SetupH& S, BlockH& B, const SurfView* F) {
  for (int i = 0; i < 3; i++) B.dx[i] = 1. / (double)(B.nx[i] - 1);
  const int ny = B.nx[1], nz = B.nx[2];
  for (int c = 0; c < S.ndim; c++)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < B.nx[0]; i++)
      for (int j = 0; j < B.nx[1]; j++)
        for (int k = 0; k < B.nx[2]; k++) {
          const double p = (double)i * B.dx[0], q = (double)j * B.dx[1], r = (double)k * B.dx[2];
          double v = ((1. - p) * F[0](c, j, k) + p * F[1](c, j, k) + (1. - q) * F[2](c, i, k) + q * F[3](c, i, k));
          if (S.ndim == 3) v += (1. - r) * F[4](c, i, j) + r * F[5](c, i, j);
          v -= ((1. - q) * (1. - p) * F[0](c, 0, k) + (1. - q) * p * F[1](c, 0, k) + q * (1. - p) * F[0](c, ny - 1, k) +
                q * p * F[1](c, ny - 1, k));
          if (S.ndim == 3) {
            v -= ((1. - p) * (1. - r) * F[0](c, j, 0) + p * (1. - r) * F[1](c, j, 0) + (1. - q) * (1. - r) * F[2](c, i, 0) +
                  q * (1. - r) * F[3](c, i, 0) + (1. - p) * r * F[0](c, j, nz - 1) + p * r * F[1](c, j, nz - 1) +
                  (1. - q) * r * F[2](c, i, nz - 1) + q * r * F[3](c, i, nz - 1));
            v += ((1. - p) * (1. - q) * (1. - r) * F[0](c, 0, 0) + p * (1. - q) * (1. - r) * F[1](c, 0, 0) +
                  (1. - p) * q * (1. - r) * F[0](c, ny - 1, 0) + (1. - p) * (1. - q) * r * F[0](c, 0, nz - 1) +
                  p * q * (1. - r) * F[1](c, ny - 1, 0) + p * (1. - q) * r * F[1](c, 0, nz - 1) +
                  (1. - p) * q * r * F[0](c, ny - 1, nz - 1) + p * q * r * F[1](c, ny - 1, nz - 1));
          }
          S.x[(size_t)c * S.N + S.idx(B.xm[0] + i, B.xm[1] + j, B.xm[2] + k)] = v;
        }
  // dX_c / dxi_d along grid lines, one-sided rows near block edges
  const int s = S.fd.s;
  const size_t nb = (size_t)B.nx[0] * B.nx[1] * B.nx[2];
  gvec J9(9 * nb);  // J9[(c*3+d)*nb + local]; entries with c, d < ndim are written below
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++)
      if (c >= S.ndim || d >= S.ndim) {
        const double v = c == d ? 1. : 0.;
        double* J = J9.data() + (size_t)(c * 3 + d) * nb;
#pragma omp parallel for schedule(static)
        for (long long q = 0; q < (long long)nb; q++) J[q] = v;
      }
  const size_t gstride[3] = {(size_t)S.nx[1] * S.nx[2], (size_t)S.nx[2], 1};
  for (int d = 0; d < S.ndim; d++)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < B.nx[0]; i++)
      for (int j = 0; j < B.nx[1]; j++)
        for (int k = 0; k < B.nx[2]; k++) {
          const int loc[3] = {i, j, k};
          const int pos = loc[d];
          const size_t g = S.idx(B.xm[0] + i, B.xm[1] + j, B.xm[2] + k);
          const size_t line0 = g - (size_t)pos * gstride[d];  // global index of this line's first point
          const size_t lidx = ((size_t)i * B.nx[1] + j) * B.nx[2] + k;
          const int edge = 2 * (s - 1), n = B.nx[d];
          for (int c = 0; c < S.ndim; c++) {
            const double* X = S.x.data() + (size_t)c * S.N + line0;
            double acc = 0.;
            if (pos < edge) {
              for (int t = 0; t < 3 * (s - 1); t++) acc += S.fd.c[pos + 1][t] * X[(size_t)t * gstride[d]] / B.dx[d];
            } else if (pos < n - edge) {
              for (int t = 0; t < 2 * s - 1; t++)
                acc += S.fd.c[0][t] * X[(size_t)(pos - s + 1 + t) * gstride[d]] / B.dx[d];
            } else {
              for (int t = 0; t < 3 * (s - 1); t++) acc -= S.fd.c[n - pos][t] * X[(size_t)(n - 1 - t) * gstride[d]] / B.dx[d];
            }
            J9[(size_t)(c * 3 + d) * nb + lidx] = acc;
          }
        }
  bool bad_jac = false;
#pragma omp parallel for schedule(static) reduction(|| : bad_jac)
  for (int i = 0; i < B.nx[0]; i++)
    for (int j = 0; j < B.nx[1]; j++)
      for (int k = 0; k < B.nx[2]; k++) {
        const size_t lidx = ((size_t)i * B.nx[1] + j) * B.nx[2] + k;
        auto X = [&](int c, int d) { return J9[(size_t)(c * 3 + d) * nb + lidx]; };
        const double jac = (X(0, 0) * (X(1, 1) * X(2, 2) - X(1, 2) * X(2, 1)) - X(1, 0) * (X(0, 1) * X(2, 2) - X(0, 2) * X(2, 1)) +
                            X(2, 0) * (X(0, 1) * X(1, 2) - X(0, 2) * X(1, 1)));
        if (!(jac > 0.)) bad_jac = true;
        const size_t g = S.idx(B.xm[0] + i, B.xm[1] + j, B.xm[2] + k);
        S.jac[g] = jac;
        for (int l = 0; l < S.ndim; l++)
          for (int m = 0; m < S.ndim; m++)
            S.metric[(size_t)(l * S.ndim + m) * S.N + g] =
                ((X((m + 1) % 3, (l + 1) % 3) * X((m + 2) % 3, (l + 2) % 3) - X((m + 1) % 3, (l + 2) % 3) * X((m + 2) % 3, (l + 1) % 3)) / jac);
      }
  if (bad_jac) fail("non-positive Jacobian in block grid (grid is not smooth enough)");
}

// unit normal and SAT scale from one metric row at point g (boundary.cpp:206-236)
double unit_normal(const SetupH& S, int d, size_t g, double* nrm) {
  double len = 0.;
  for (int k = 0; k < S.ndim; k++) {
    const double m = S.metric_at(d * S.ndim + k, g);
    len += m * m;
    nrm[k] = m;
  }
  len = sqrt(len);
  for (int k = 0; k < S.ndim; k++) nrm[k] /= len;
  return len;
}

bool edges_match(const SetupH& S, const SurfView& s1, int e1, const SurfView& s2, int e2) {
  // edges: 0 second index 0, 1 first index 0, 2 second index last, 3 first index last
  auto at = [](const SurfView& v, int e, int c, int m) {
    if (e == 0) return v(c, m, 0);
    if (e == 1) return v(c, 0, m);
    if (e == 2) return v(c, m, v.n1 - 1);
    return v(c, v.n0 - 1, m);
  };
  const int l1 = (e1 & 1) ? s1.n1 : s1.n0, l2 = (e2 & 1) ? s2.n1 : s2.n0;
  if (l1 != l2) return false;
  for (int c = 0; c < S.ndim; c++)
    for (int m = 0; m < l1; m++) {
      const double a = at(s1, e1, c, m), b = at(s2, e2, c, m);
      if (fabs(a - b) > 1.e-8 + 1.e-5 * fabs(b)) return false;
    }
  return true;
}

void read_stack(const std::string& path, int k, size_t n, std::vector<double>* out, const std::string& open_msg,
                const char* what) {
  std::vector<double> all = load_f8(path, (size_t)k * n, open_msg, what);
  for (int q = 0; q < k; q++) out[q].assign(all.begin() + (size_t)q * n, all.begin() + (size_t)(q + 1) * n);
}

double min_dx_over_cs(const SetupH& S) {  // block.cpp:368-403, domain.cpp:199-226
  double best = 0.;
  for (const BlockH& B : S.blocks) {
    double worst = 0.;
    if (S.gb) {
      worst = S.gb->block_worst(S, B);
      const double v = 1. / worst;
      if (v > 1.e-14 && (best <= 1.e-14 || v < best)) best = v;
      continue;
    }
#pragma omp parallel for schedule(static) reduction(max : worst)
    for (int i = B.mlb[0]; i < B.prb[0]; i++)
      for (int j = B.mlb[1]; j < B.prb[1]; j++)
        for (int k = B.mlb[2]; k < B.prb[2]; k++) {
          const size_t g = S.idx(i, j, k);
          for (int l = 0; l < S.ndim; l++) {
            double t = 0.;
            for (int m = 0; m < S.ndim; m++) {
              const double q = S.metric[(size_t)(l * S.ndim + m) * S.N + g];
              t += q * q;
            }
            const double cs = S.hetmat ? sqrt(S.mat[2 * S.N + g] / S.mat[g]) : B.cs;
            t = cs * sqrt(t) / B.dx[l];
            if (t > worst) worst = t;
          }
        }
    const double v = 1. / worst;
    if (v > 1.e-14 && (best <= 1.e-14 || v < best)) best = v;
  }
  if (best <= 1.e-14) fail("Error in domain.cpp get_min_dx -- a process does not have any grid spacing values");
  return best;
}

}  // namespace

void SetupH::gather(const std::vector<size_t>& idx) {
  if (!gb) return;
  std::vector<size_t> todo;
  for (size_t g : idx)
    if (gslot.emplace(g, gslot.size()).second) todo.push_back(g);
  if (todo.empty()) return;
  const size_t base = gj.size();
  gx.resize((base + todo.size()) * ndim);
  gm.resize((base + todo.size()) * ndim * ndim);
  gj.resize(base + todo.size());
  gb->gather(*this, todo, gx.data() + base * ndim, gm.data() + base * ndim * ndim, gj.data() + base);
}

bool block_metric_uniform(const SetupH& S, const BlockH& B, double mdiag[3]) {
  if (S.gb) fail("block_metric_uniform: the grid arrays are on the device");
  const int nd = S.ndim;
  const size_t N = S.N;
  const size_t gc = S.idx((B.mlb[0] + B.prb[0]) / 2, (B.mlb[1] + B.prb[1]) / 2, (B.mlb[2] + B.prb[2]) / 2);
  const double jc = S.jac[gc];
  for (int d = 0; d < nd; d++) mdiag[d] = S.metric[(size_t)(d * nd + d) * N + gc];
  bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
  for (int i = B.mlb[0]; i < B.prb[0]; i++)
    for (int j = B.mlb[1]; j < B.prb[1]; j++)
      for (int k = B.mlb[2]; k < B.prb[2]; k++) {
        const size_t g = S.idx(i, j, k);
        if (fabs(S.jac[g] - jc) > 1e-9 * fabs(jc)) bad = true;
        for (int l = 0; l < nd; l++)
          for (int m = 0; m < nd; m++) {
            const double v = S.metric[(size_t)(l * nd + m) * N + g];
            if (l == m ? fabs(v - mdiag[l]) > 1e-9 * fabs(mdiag[l]) : fabs(v) > 1e-9 * fabs(mdiag[l])) bad = true;
          }
      }
  return !bad;
}

void compile_problem(const std::string& path, SetupH& S, bool outputs, bool quiet, GridBuilder* gb) {
  S.gb = gb;
  Scanner sc(path);
  // [fdfault.problem], problem.cpp:27-44
  if (!sc.seek("[fdfault.problem]")) fail("Error reading problem from input file");
  S.name = sc.word();
  S.datadir = sc.word();
  S.nt = sc.integer();
  S.dt = sc.real();
  S.ttot = sc.real();
  S.cfl = sc.real();
  S.ninfo = sc.integer();
  S.rkorder = sc.integer();
  init_rk(S.rk, S.rkorder);
  // [fdfault.domain], domain.cpp:47-70
  if (!sc.seek("[fdfault.domain]")) fail("Error reading domain from input file");
  S.ndim = sc.integer();
  S.mode = sc.integer();
  for (int i = 0; i < 3; i++) S.nx[i] = sc.integer();
  for (int i = 0; i < 3; i++) S.nb[i] = sc.integer();
  std::vector<int> nxb[3];
  for (int i = 0; i < 3; i++) {
    if (S.nb[i] <= 0) fail("invalid number of blocks");
    for (int j = 0; j < S.nb[i]; j++) nxb[i].push_back(sc.integer());
  }
  const int nif = sc.integer();
  std::vector<std::string> iftype;
  for (int i = 0; i < nif; i++) iftype.push_back(sc.word());
  S.sbporder = sc.integer();
  const std::string material = sc.word();
  if (S.ndim != 2 && S.ndim != 3) fail("ndim must be 2 or 3");
  if (S.mode != 2 && S.mode != 3) fail("mode must be 2 or 3");
  if (material != "elastic" && material != "plastic") fail("material must be elastic or plastic");
  if (S.ndim == 2 && S.nx[2] != 1) fail("2d problems require nz = 1");
  S.plastic = (material == "plastic");
  for (int i = 0; i < 3; i++) {
    int sum = 0;
    for (int v : nxb[i]) sum += v;
    if (sum != S.nx[i]) fail("block sizes do not sum to grid size");
  }
  init_sbp(S.fd, S.sbporder);
  // [fdfault.fields], fields.cpp:36-41 and 53-164
  if (!sc.seek("[fdfault.fields]")) fail("Error reading fields from input file");
  for (int i = 0; i < 6; i++) S.s0[i] = sc.real();
  const std::string loadfile = sc.word(), matfile = sc.word();
  S.plastic_tensor = sc.integer() != 0;
  if (S.ndim == 3) { S.nf = 9; S.nmat = 3; S.ns = 6; S.nv = 3; for (int i = 0; i < 6; i++) S.sidx[i] = i; }
  else if (S.mode == 2) { S.nf = 5; S.nmat = 3; S.ns = 3; S.nv = 2; S.sidx[0] = 0; S.sidx[1] = 1; S.sidx[2] = 3; }
  else { S.nf = 3; S.nmat = 2; S.ns = 2; S.nv = 1; S.sidx[0] = 2; S.sidx[1] = 4; }
  if (S.plastic) {
    S.nfp = S.plastic_tensor ? 8 : 2;
    if (S.ndim == 2 && S.mode == 2) { S.nf = 6; S.ns = 4; S.sidx[3] = 5; }
  }
  S.N = (size_t)S.nx[0] * S.nx[1] * S.nx[2];
  // (blocks tile the grid, so build_block_grid writes every entry)
  if (!S.gb) {
    S.x.resize((size_t)S.ndim * S.N);
    S.metric.resize((size_t)S.ndim * S.ndim * S.N);
    S.jac.resize(S.N);
  }
  if (loadfile != "none") {
    S.hetstress = true;
    S.s = load_f8(loadfile, (size_t)S.ns * S.N, "Error opening stress file in fields.cpp", "heterogeneous stress");
  }
  if (matfile != "none") {
    S.hetmat = true;
    S.mat = load_f8(matfile, (size_t)S.nmat * S.N, "Error opening material file in fields.cpp", "heterogeneous material");
  }
  if (S.hetmat && S.ndim == 2 && S.mode == 3)
    fail("heterogeneous material in mode III is not supported (the reference reads past its material array there)");
  // [fdfault.operator] (optional), block.cpp:91-107
  S.cdiss = sc.seek("[fdfault.operator]") ? sc.real() : 0.;
  if (!(S.cdiss > 0.)) S.cdiss = 0.;
  // blocks, block.cpp:22-264
  const int s = S.sbporder;
  S.blocks.resize((size_t)S.nb[0] * S.nb[1] * S.nb[2]);
  int start[3][64];
  for (int d = 0; d < 3; d++) {
    if (S.nb[d] > 64) fail("too many blocks");
    for (int j = 0; j < S.nb[d]; j++) start[d][j] = j ? start[d][j - 1] + nxb[d][j - 1] : 0;
  }
  for (int bi = 0; bi < S.nb[0]; bi++)
    for (int bj = 0; bj < S.nb[1]; bj++)
      for (int bk = 0; bk < S.nb[2]; bk++) {
        BlockH& B = S.blocks[((size_t)bi * S.nb[1] + bj) * S.nb[2] + bk];
        const int cc[3] = {bi, bj, bk};
        const std::string tag = std::to_string(bi) + std::to_string(bj) + std::to_string(bk);
        if (!sc.seek("[fdfault.block" + tag + "]")) fail("Error reading block " + tag + " from input file");
        for (int d = 0; d < 3; d++) {
          B.coords[d] = cc[d];
          B.nx[d] = nxb[d][cc[d]];
          B.xm[d] = start[d][cc[d]];
        }
        B.rho = sc.real();
        B.lam = sc.real();
        B.G = sc.real();
        if (S.plastic) {
          B.pmu = sc.real();
          B.pc = sc.real();
          B.pbeta = sc.real();
          B.peta = sc.real();
        }
        for (int d = 0; d < S.ndim; d++) B.x0[d] = sc.real();
        for (int d = 0; d < S.ndim; d++) B.L[d] = sc.real();
        for (int f = 0; f < 2 * S.ndim; f++) B.btype[f] = sc.word();
        for (int f = 0; f < 2 * S.ndim; f++) B.bfile[f] = sc.word();
        for (int d = 0; d < S.ndim; d++)
          if (B.nx[d] < 4 * s - 3) fail("block too small for the finite difference method");
        for (int d = 0; d < 3; d++) {  // block.cpp:165-179 with one process
          B.mlb[d] = B.xm[d];
          const bool line = B.nx[d] > 1;
          B.mc[d] = line ? B.mlb[d] + 2 * (s - 1) : B.mlb[d];
          B.mrb[d] = line ? B.mlb[d] + B.nx[d] - 2 * (s - 1) : B.mlb[d] + B.nx[d];
          B.prb[d] = line ? B.mrb[d] + 2 * (s - 1) : B.mrb[d];
        }
        B.cs = sqrt(B.G / B.rho);  // material.cpp:122-144
        B.cp = sqrt((B.lam + 2. * B.G) / B.rho);
        B.zs = B.rho * B.cs;
        B.zp = B.rho * B.cp;
        std::vector<double> surf[6];
        SurfView view[6];
        for (int f = 0; f < 2 * S.ndim; f++) {
          int n0, n1;
          surf[f] = face_surface(S, B, f, n0, n1);
          view[f] = SurfView{surf[f].data(), n0, n1};
        }
        static const int s1[12] = {0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 3, 3}, s2[12] = {2, 3, 2, 3, 4, 5, 4, 5, 4, 5, 4, 5};
        static const int e1[12] = {1, 3, 1, 3, 0, 2, 0, 2, 0, 2, 0, 2}, e2[12] = {1, 1, 3, 3, 1, 1, 3, 3, 0, 0, 2, 2};
        for (int e = 0; e < (S.ndim == 3 ? 12 : 4); e++)
          if (!edges_match(S, view[s1[e]], e1[e], view[s2[e]], e2[e])) fail("Surface edges do not match in block.cpp");
        if (S.gb) {
          for (int i = 0; i < 3; i++) B.dx[i] = 1. / (double)(B.nx[i] - 1);
          int n0s[6], n1s[6];
          for (int f = 0; f < 2 * S.ndim; f++) {
            n0s[f] = view[f].n0;
            n1s[f] = view[f].n1;
          }
          S.gb->build_block(S, B, surf, n0s, n1s);
        } else {
          build_block_grid(S, B, view);
        }
      }
  if (S.gb) {
    if (!S.gb->jacobians_positive()) fail("non-positive Jacobian in block grid (grid is not smooth enough)");
    // the points whose grid values the host needs: every face point of
    // every block (normals and SAT scales) and each block's centre point
    std::vector<size_t> need;
    for (const BlockH& B : S.blocks) {
      need.push_back(S.idx((B.mlb[0] + B.prb[0]) / 2, (B.mlb[1] + B.prb[1]) / 2, (B.mlb[2] + B.prb[2]) / 2));
      for (int f = 0; f < 2 * S.ndim; f++) {
        const int d = f / 2, a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
        int c[3];
        c[d] = B.xm[d] + ((f & 1) ? B.nx[d] - 1 : 0);
        for (int u = 0; u < B.nx[a]; u++)
          for (int v = 0; v < B.nx[b]; v++) {
            c[a] = B.xm[a] + u;
            c[b] = B.xm[b] + v;
            need.push_back(S.idx(c[0], c[1], c[2]));
          }
      }
    }
    S.gather(need);
  }
  // external boundary faces, boundary.cpp:15-240
  for (BlockH& B : S.blocks)
    for (int f = 0; f < 2 * S.ndim; f++) {
      FaceBC& F = B.face[f];
      const std::string& t = B.btype[f];
      if (t != "none" && t != "absorbing" && t != "free" && t != "rigid") fail("unknown boundary type " + t);
      if (t == "none") continue;
      F.active = true;
      if (t == "absorbing") { F.r = 0.; F.alpha = 2.; F.theta = 0.; }
      else if (t == "free") { F.r = -1.; F.alpha = 1.; F.theta = 0.; }
      else { F.r = 1.; F.alpha = 0.; F.theta = 1.; }
      F.cp = B.cp; F.cs = B.cs; F.zp = B.zp; F.zs = B.zs;
      const double q = F.cs / F.cp;
      F.gamma = 1. - 2. * (q * q);
      const int d = f / 2, a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
      F.n0 = B.nx[a];
      F.n1 = B.nx[b];
      const size_t n = (size_t)F.n0 * F.n1;
      F.nrm.assign((size_t)S.ndim * n, 0.);
      F.dl.assign(n, 0.);
      for (int u = 0; u < F.n0; u++)
        for (int v = 0; v < F.n1; v++) {
          int c[3];
          c[d] = B.xm[d] + ((f & 1) ? B.nx[d] - 1 : 0);
          c[a] = B.xm[a] + u;
          c[b] = B.xm[b] + v;
          double nn[3];
          const double len = unit_normal(S, d, S.idx(c[0], c[1], c[2]), nn);
          const size_t o = (size_t)u * F.n1 + v;
          for (int k = 0; k < S.ndim; k++) F.nrm[(size_t)k * n + o] = (f & 1) ? nn[k] : -nn[k];
          F.dl[o] = len / (S.fd.h0 * B.dx[d]);
        }
    }
  // interfaces, interface.cpp:17-534 and the friction-law constructors
  S.ifaces.resize(nif);
  for (int ii = 0; ii < nif; ii++) {
    IfaceH& I = S.ifaces[ii];
    const std::string& t = iftype[ii];
    if (t == "locked") I.law = LAW_LOCKED;
    else if (t == "frictionless") I.law = LAW_FRICTIONLESS;
    else if (t == "slipweak") I.law = LAW_SLIPWEAK;
    else if (t == "stz") I.law = LAW_STZ;
    else fail("unknown interface type " + t);
    const std::string hdr = "[fdfault.interface" + std::to_string(ii) + "]";
    if (!sc.seek(hdr)) fail("Error reading interface " + std::to_string(ii) + " from input file");
    const std::string dir = sc.word();
    int c1[3], c2[3];
    for (int i = 0; i < 3; i++) c1[i] = sc.integer();
    for (int i = 0; i < 3; i++) c2[i] = sc.integer();
    if (dir == "x") I.dir = 0;
    else if (dir == "y") I.dir = 1;
    else if (dir == "z" && S.ndim == 3) I.dir = 2;
    else fail("bad interface direction");
    const int d = I.dir, a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
    for (int i = 0; i < 3; i++) {
      if (c2[i] != c1[i] + (i == d ? 1 : 0)) fail("interface blocks are not adjacent");
      if (c1[i] < 0 || c2[i] >= S.nb[i]) fail("interface block index out of range");
    }
    I.b1 = (c1[0] * S.nb[1] + c1[1]) * S.nb[2] + c1[2];
    I.b2 = (c2[0] * S.nb[1] + c2[1]) * S.nb[2] + c2[2];
    const BlockH& B1 = S.blocks[I.b1];
    const BlockH& B2 = S.blocks[I.b2];
    if (B1.nx[a] != B2.nx[a] || B1.nx[b] != B2.nx[b]) fail("interface blocks do not match in size");
    I.n0 = B1.nx[a];
    I.n1 = B1.nx[b];
    for (int i = 0; i < 3; i++) {
      I.org[i] = (i == d) ? B1.xm[i] + B1.nx[i] - 1 : B1.xm[i];
      I.xm[i] = B2.xm[i];
      I.xp[i] = (i == d) ? B2.xm[i] : B2.xm[i] + B2.nx[i] - 1;
    }
    const size_t n = (size_t)I.n0 * I.n1;
    I.nrm.assign((size_t)S.ndim * n, 0.);
    I.dl1.assign(n, 0.);
    I.dl2.assign(n, 0.);
    for (int u = 0; u < I.n0; u++)
      for (int v = 0; v < I.n1; v++) {
        int c[3];
        c[d] = I.org[d];
        c[a] = I.org[a] + u;
        c[b] = I.org[b] + v;
        const size_t g1 = S.idx(c[0], c[1], c[2]);
        c[d] += 1;
        const size_t g2 = S.idx(c[0], c[1], c[2]);
        for (int k = 0; k < S.ndim; k++) {  // interface.cpp:333-347 edge agreement
          const double xa = S.x_at(k, g1), xb = S.x_at(k, g2);
          if (!(fabs(xa - xb) <= 1.e-8 + 1.e-5 * fabs(xb))) fail("interface block edges do not match");
        }
        double nn[3], nn2[3];
        const double l1 = unit_normal(S, d, g1, nn);
        const double l2 = unit_normal(S, d, g2, nn2);
        const size_t o = (size_t)u * I.n1 + v;
        for (int k = 0; k < S.ndim; k++) I.nrm[(size_t)k * n + o] = nn[k];
        I.dl1[o] = l1 / (S.fd.h0 * B1.dx[d]);
        I.dl2[o] = l2 / (S.fd.h0 * B2.dx[d]);
      }
    I.cp1 = B1.cp; I.cs1 = B1.cs; I.zp1 = B1.zp; I.zs1 = B1.zs;
    I.cp2 = B2.cp; I.cs2 = B2.cs; I.zp2 = B2.zp; I.zs2 = B2.zs;
    const double q1 = I.cs1 / I.cp1, q2 = I.cs2 / I.cp2;
    I.gamma1 = 1. - 2. * (q1 * q1);
    I.gamma2 = 1. - 2. * (q2 * q2);
    if (I.law == LAW_LOCKED) continue;
    // perturbation coordinates come from block 1's box (friction.cpp:159-183)
    const double x2[2] = {B1.x0[a], B1.x0[b]}, l2[2] = {B1.L[a], B1.L[b]};
    sc.seek(hdr);
    if (!sc.seek("[fdfault.friction]", false)) fail("Error reading interface " + std::to_string(ii) + " from input file");
    const int nloads = sc.integer();
    for (int k = 0; k < nloads; k++) {
      PertH p = make_pert(sc, 3, I, x2, l2);
      if (S.ndim == 2 && S.mode == 2) p.amp[2] = 0.;
      else if (S.ndim == 2 && S.mode == 3) p.amp[1] = 0.;
      I.loads.push_back(std::move(p));
    }
    const std::string lf = sc.word();
    if (lf != "none") {
      I.load_file = true;
      read_stack(lf, 3, n, I.lf, "Error opening load file in friction.cpp", "load");
      if (S.ndim == 2 && S.mode == 2) std::fill(I.lf[2].begin(), I.lf[2].end(), 0.);
      else if (S.ndim == 2 && S.mode == 3) std::fill(I.lf[1].begin(), I.lf[1].end(), 0.);
    }
    if (I.law == LAW_SLIPWEAK) {
      sc.seek(hdr);
      if (!sc.seek("[fdfault.slipweak]", false)) fail("Error reading slipweak from input file");
      const int np = sc.integer();
      for (int k = 0; k < np; k++) I.perts.push_back(make_pert(sc, 6, I, x2, l2));
      const std::string pf = sc.word();
      if (pf != "none") { I.param_file = true; read_stack(pf, 6, n, I.pf, "Error opening file in friction.cpp", "slip weakening parameters"); }
    } else if (I.law == LAW_STZ) {
      sc.seek(hdr);
      if (!sc.seek("[fdfault.stz]", false)) fail("Error reading stz from input file");
      const double chi0 = sc.real();
      const std::string sfile = sc.word();
      const int np = sc.integer();
      for (int k = 0; k < np; k++) I.perts.push_back(make_pert(sc, 9, I, x2, l2));
      const std::string pf = sc.word();
      I.state0.assign(n, 0.);
      if (sfile != "none") I.state0 = load_f8(sfile, n, "Error opening state file in friction.cpp", "state");
      for (double& v : I.state0) v += chi0;
      if (pf != "none") { I.param_file = true; read_stack(pf, 9, n, I.pf, "Error opening file in stz.cpp", "stz parameters"); }
    }
  }
  // time step, problem.cpp:94-134
  const double dxm = min_dx_over_cs(S);
  if ((S.ttot > 0. && S.nt > 0) && (S.dt == 0 && S.cfl == 0.)) {
    S.dt = S.ttot / (double)S.nt;
    S.cfl = S.dt / dxm;
  } else {
    if (S.dt > 0.) {
      if (S.cfl > 0. && !quiet) printf("Cannot specify both dt and cfl, defaulting to dt\n");
      S.cfl = S.dt / dxm;
    } else {
      S.dt = S.cfl * dxm;
    }
    if (S.ttot > 0.) {
      if (S.nt > 0 && !quiet) printf("Cannot specify both ttot and nt with one of cfl or dt, defaulting to ttot\n");
      S.nt = (int)ceil(S.ttot / S.dt);
      S.ttot = (double)S.nt * S.dt;
    } else {
      S.ttot = (double)S.nt * S.dt;
    }
  }
  if (S.cfl > 1. && !quiet) printf("Warning: CFL ratio > 1, numerical instability is likely\n");
  // per-direction stencil constants with the y/z rescaling of block.cpp:1166-1169,1883-1886
  for (BlockH& B : S.blocks) {
    B.K[0][0] = S.dt / B.rho / B.dx[0];
    B.K[0][1] = S.dt * (2. * B.G + B.lam) / B.dx[0];
    B.K[0][2] = S.dt * B.G / B.dx[0];
    B.K[0][3] = S.dt * B.lam / B.dx[0];
    for (int d = 1; d < 3; d++) {
      const double r = (d < S.ndim) ? B.dx[d - 1] / B.dx[d] : 0.;
      for (int q = 0; q < 4; q++) B.K[d][q] = B.K[d - 1][q] * r;
    }
  }
  // output list (outputlist.cpp:40-66, outputunit.cpp:13-343) and front list
  if (outputs) {
    if (!sc.seek("[fdfault.outputlist]")) fail("Error reading outputlist from input file");
    std::string line;
    while (sc.line(line)) {
      if (line.empty()) break;
      OutputH U;
      U.name = line;
      U.field = sc.word();
      int tm = sc.integer(), tp = sc.integer(), ts = sc.integer();
      int xm[3], xp[3], xs[3];
      for (int i = 0; i < 3; i++) { xm[i] = sc.integer(); xp[i] = sc.integer(); xs[i] = sc.integer(); }
      sc.line(line);
      if (ts <= 0 || tp < tm) fail("invalid output time range");
      for (int i = 0; i < 3; i++)
        if (xm[i] >= S.nx[i] || xs[i] <= 0 || xp[i] < xm[i]) fail("invalid output index range");
      static const char* g3[] = {"vx", "vy", "vz", "sxx", "sxy", "sxz", "syy", "syz", "szz",
                                 "lambda", "gammap", "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
      static const char* g2[] = {"vx", "vy", "sxx", "sxy", "syy", "szz", "lambda", "gammap",
                                 "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
      static const char* g2m3[] = {"vz", "sxz", "syz", "lambda", "gammap", "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
      const char** gl = (S.ndim == 3) ? g3 : (S.mode == 2 ? g2 : g2m3);
      const int ng = (S.ndim == 3) ? 17 : (S.mode == 2 ? 14 : 11);
      int gidx = -1;
      for (int i = 0; i < ng; i++) if (U.field == gl[i]) gidx = i;
      const std::string& fn = U.field;
      bool ifield = false;
      int ifx = -1;
      if (gidx < 0) {
        // interface quantities; vector components exist only in-plane
        const char* names[] = {"U", "V", "S", "Sn", "Ux", "Vx", "Sx", "state", "Uy", "Vy", "Sy", "Uz", "Vz", "Sz"};
        const int ids[] = {0, 1, 2, 3, 4, 5, 6, 7, 4, 5, 6, 4, 5, 6};
        for (int i = 0; i < 14; i++) if (fn == names[i]) { ifx = ids[i]; ifield = true; }
        if (!ifield) fail("Error in specifying output field in outputunit.cpp");
        if (S.ndim == 2 && S.mode == 2 && fn.size() == 2 && fn[1] == 'z') fail("Error in specifying output field in outputunit.cpp");
        if (S.ndim == 2 && S.mode == 3 && fn.size() == 2 && (fn[1] == 'x' || fn[1] == 'y'))
          fail("Error in specifying output field in outputunit.cpp");
      } else {
        if ((fn == "lambda" || fn == "gammap") && !S.plastic) fail("plastic output requested for elastic problem");
        if (fn.size() == 4 && fn.compare(0, 2, "ep") == 0 && !S.plastic_tensor) fail("plastic strain output requires plastic tensor");
        if (S.ndim == 2 && S.mode == 2 && fn == "szz" && !S.plastic) fail("szz output requires plasticity in mode 2");
      }
      U.fidx = ifield ? ifx : gidx;
      if (ifield) {
        for (int i = 0; i < nif; i++) {
          const IfaceH& I = S.ifaces[i];
          const int d = I.dir;
          if (xm[d] != xp[d]) continue;
          bool ok = (xm[d] == I.xm[d] || xm[d] == I.xm[d] - 1);
          for (int q = 0; q < 3; q++)
            if (q != d) ok = ok && xm[q] >= I.xm[q] && xm[q] <= I.xp[q];
          if (ok) { U.iface = i; break; }
        }
        if (U.iface < 0) fail("Could not find interface corresponding to input indices in outputitem.cpp");
        const IfaceH& I = S.ifaces[U.iface];
        if (I.law == LAW_LOCKED) fail("Interface for output is not frictional in outputunit.cpp");
        if (fn == "state" && I.law != LAW_STZ) fail("Interface for output does not have a state variable in outputunit.cpp");
        if (fn.size() == 2 && fn != "Sn") {
          const int axis = (S.ndim == 2 && S.mode == 3) ? 2 : fn[1] - 'x';
          if (axis == I.dir) fail("Error in specifying output field in outputunit.cpp");
          U.comp = (S.ndim == 3 && axis != ((I.dir == 0) ? 1 : 0)) ? 1 : 0;
        }
      }
      U.tm = tm;
      U.tp = (tp > S.nt) ? S.nt : tp;
      U.tp -= (U.tp - U.tm) % ts;
      U.ts = ts;
      for (int i = 0; i < 3; i++) {
        if (U.iface < 0) {
          U.xm[i] = xm[i] < 0 ? 0 : xm[i];
          U.xp[i] = (xp[i] >= S.nx[i]) ? S.nx[i] - 1 : xp[i] - (xp[i] - xm[i]) % xs[i];
        } else {
          const IfaceH& I = S.ifaces[U.iface];
          U.xm[i] = xm[i] < I.xm[i] ? I.xm[i] : xm[i];
          U.xp[i] = (xp[i] > I.xp[i]) ? I.xp[i] : xp[i] - (xp[i] - xm[i]) % xs[i];
          if (U.xm[i] > U.xp[i]) U.xp[i] = U.xm[i];
        }
        U.xs[i] = xs[i];
        U.nx[i] = (U.xp[i] - U.xm[i]) / U.xs[i] + 1;
      }
      if (S.ndim == 2) { U.xm[2] = 0; U.xp[2] = 0; U.xs[2] = 1; U.nx[2] = 1; }
      for (int i = 0; i < 3; i++) if (U.nx[i] == 1) U.xs[i] = 1;
      S.outputs.push_back(U);
    }
  }
  if (!sc.seek("[fdfault.frontlist]")) fail("Error reading frontlist from input file");
  S.has_front = sc.integer() != 0;
  if (S.has_front) {
    const std::string ff = sc.word();
    S.front_value = sc.real();
    if (!(S.front_value > 0.) || (ff != "V" && ff != "U")) fail("invalid front parameters");
    S.front_field = (ff == "V") ? 0 : 1;
  }
}

}  // namespace fdf
