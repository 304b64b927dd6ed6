// fdfault CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Single-rank FP64 restatement of the reference solver.  Names of routines
// follow the reference's classes so a reader can hold the two side by side;
// every routine cites the reference file:line whose arithmetic it restates.
// Compiled with -O2 -ffp-contract=off so no FMA contraction changes rounding.

#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

struct Err : std::runtime_error {
  using std::runtime_error::runtime_error;
};

static thread_local std::string g_err;

// ---------------------------------------------------------------------------
// input text: emulates the std::getline / operator>> scanning the reference
// constructors perform on the .in file (e.g. domain.cpp:33-71).
class InText {
 public:
  explicit InText(const std::string& path) {
    std::ifstream in(path, std::ios::in | std::ios::binary);
    if (!in.is_open()) throw Err("Error opening input file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    text_ = ss.str();
  }
  void rewind() { pos_ = 0; eof_ = false; }
  bool getline(std::string& line) {
    if (pos_ >= text_.size()) { eof_ = true; return false; }
    size_t e = text_.find('\n', pos_);
    if (e == std::string::npos) {
      line = text_.substr(pos_);
      pos_ = text_.size();
      eof_ = true;
      return true;
    }
    line = text_.substr(pos_, e - pos_);
    pos_ = e + 1;
    return true;
  }
  // scan forward to a header line; false if the scan hit end of file
  bool scan(const std::string& header) {
    std::string line;
    while (getline(line)) {
      if (line == header) break;
    }
    return !eof_;
  }
  bool at_eof() const { return eof_; }
  std::string tok() {
    while (pos_ < text_.size() && isspace((unsigned char)text_[pos_])) pos_++;
    if (pos_ >= text_.size()) { eof_ = true; throw Err("unexpected end of input file"); }
    size_t s = pos_;
    while (pos_ < text_.size() && !isspace((unsigned char)text_[pos_])) pos_++;
    if (pos_ >= text_.size()) eof_ = true;
    return text_.substr(s, pos_ - s);
  }
  double dbl() {
    std::string t = tok();
    char* end = nullptr;
    double v = strtod(t.c_str(), &end);
    if (end == t.c_str()) throw Err("bad number '" + t + "' in input file");
    return v;
  }
  int integer() {
    std::string t = tok();
    char* end = nullptr;
    long v = strtol(t.c_str(), &end, 10);
    if (end == t.c_str()) throw Err("bad integer '" + t + "' in input file");
    return (int)v;
  }

 private:
  std::string text_;
  size_t pos_ = 0;
  bool eof_ = false;
};

// open_msg: the reference's message for a file it cannot open
// (fields.cpp:497,627, friction.cpp:555,627, slipweak.cpp:358, stz.cpp:406,
// surface.cpp:44)
static std::vector<double> read_binary(const std::string& path, size_t n, const std::string& open_msg,
                                       const char* what) {
  std::vector<double> v(n);
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in.is_open()) throw Err(open_msg);
  if (!in.read((char*)v.data(), sizeof(double) * n))
    throw Err(std::string("Error reading ") + what + " from file " + path);
  return v;
}

// ---------------------------------------------------------------------------
// SBP coefficients, fd.cpp:33-155 (sbporder 2,3,4; 5/6 are unreachable there)
struct FD {
  int s = 0;
  double h0 = 0.;
  double c[7][9] = {};
  double d[7][9] = {};
};

static void fd_init(FD& fd, int order) {
  if (order < 2 || order > 4) throw Err("Error in initializing FD method");
  fd = FD();
  fd.s = order;
  auto row = [](double* r, std::initializer_list<double> v) {
    int i = 0;
    for (double x : v) r[i++] = x;
  };
  if (order == 2) {  // fd.cpp:52-63
    row(fd.c[0], {-0.5, 0., 0.5});
    row(fd.d[0], {0.5, -1., 0.5});
    row(fd.c[1], {-1., 1., 0.});
    row(fd.c[2], {-0.5, 0., 0.5});
    row(fd.d[1], {-1., 1., 0.});
    row(fd.d[2], {0.5, -1., 0.5});
    fd.h0 = 0.5;
  } else if (order == 3) {  // fd.cpp:64-79
    row(fd.c[0], {1. / 12., -2. / 3., 0., 2. / 3., -1. / 12.});
    row(fd.d[0], {-1. / 12., 1. / 3., -1. / 2., 1. / 3., -1. / 12.});
    row(fd.c[1], {-24. / 17., 59. / 34., -4. / 17., -3. / 34., 0., 0.});
    row(fd.c[2], {-0.5, 0., 0.5, 0., 0., 0.});
    row(fd.c[3], {4. / 43., -59. / 86., 0., 59. / 86., -4. / 43., 0.});
    row(fd.c[4], {3. / 98., 0., -59. / 98., 0., 32. / 49., -4. / 49.});
    row(fd.d[1], {-8. / 17., 16. / 17., -8. / 17., 0., 0., 0.});
    row(fd.d[2], {16. / 59., -36. / 59., 24. / 59., -4. / 59., 0., 0.});
    row(fd.d[3], {-8. / 43., 24. / 43., -28. / 43., 16. / 43., -4. / 43, 0.});
    row(fd.d[4], {0., -4. / 49., 16. / 49., -24. / 49., 16. / 49., -4. / 49.});
    fd.h0 = 17. / 48.;
  } else {  // fd.cpp:80-121 (the "fdmap" closure with dissipation)
    row(fd.c[0], {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.});
    row(fd.d[0], {1. / 60., -1. / 10., 1. / 4., -1. / 3., 1. / 4., -1. / 10., 1. / 60.});
    row(fd.c[1], {-1.582533518939116, 2.033378678700676, -0.141512858744873, -0.450398306578272,
                  0.104488069284042, 0.036577936277544, 0., 0., 0.});
    row(fd.c[2], {-0.462059195631158, 0., 0.287258622978251, 0.258816087376832, -0.069112065532624,
                  -0.014903449191300, 0., 0., 0.});
    row(fd.c[3], {0.071247104721830, -0.636451095137907, 0., 0.606235523609147, -0.022902190275815,
                  -0.018129342917256, 0., 0., 0.});
    row(fd.c[4], {0.114713313798970, -0.290087484386815, -0.306681191361148, 0., 0.520262285050482,
                  -0.051642265516119, 0.013435342414630, 0., 0.});
    row(fd.c[5], {-0.036210680656541, 0.105400944933782, 0.015764336127392, -0.707905442575989, 0.,
                  0.769199413962647, -0.164529643265203, 0.018281071473911, 0.});
    row(fd.c[6], {-0.011398193015050, 0.020437334208704, 0.011220896474665, 0.063183694641876,
                  -0.691649024426814, 0., 0.739709139060752, -0.147941827812150, 0.016437980868017});
    row(fd.d[1], {-0.105502234595941, 0.316506703787823, -0.316506703787823, 0.105502234595941, 0., 0., 0.,
                  0., 0.});
    row(fd.d[2], {0.071922084408557, -0.227753267293765, 0.251727295429951, -0.107883126612836,
                  0.011987014068093, 0., 0., 0., 0.});
    row(fd.d[3], {-0.159350793065290, 0.557727775728513, -0.743637034304685, 0.478052379195869,
                  -0.159350793065290, 0.026558465510882, 0., 0., 0.});
    row(fd.d[4], {0.026870684829259, -0.120918081731666, 0.241836163463333, -0.282142190707221,
                  0.201530136219444, -0.080612054487778, 0.013435342414630, 0., 0.});
    row(fd.d[5], {0., 0.018281071473911, -0.109686428843468, 0.274216072108671, -0.365621429478228,
                  0.274216072108671, -0.109686428843468, 0.018281071473911, 0.});
    row(fd.d[6], {0., 0., 0.016437980868017, -0.098627885208100, 0.246569713020251, -0.328759617360334,
                  0.246569713020251, -0.098627885208100, 0.016437980868017});
    fd.h0 = 13649. / 43200.;
  }
}

// low-storage RK coefficients, rk.cpp:8-98
struct RK {
  int order = 0, nstages = 0;
  double A[5] = {}, B[5] = {}, C[6] = {};
};

static void rk_init(RK& rk, int order) {
  rk = RK();
  rk.order = order;
  switch (order) {
    case 1:
      rk.nstages = 1;
      rk.A[0] = 0.; rk.B[0] = 1.; rk.C[0] = 0.; rk.C[1] = 1.;
      break;
    case 2:
      rk.nstages = 2;
      rk.A[0] = 0.; rk.A[1] = -1.; rk.B[0] = 1.; rk.B[1] = 0.5;
      rk.C[0] = 0.; rk.C[1] = 1.; rk.C[2] = 1.;
      break;
    case 3:
      rk.nstages = 3;
      rk.A[0] = 0.; rk.A[1] = -5. / 9.; rk.A[2] = -153. / 128.;
      rk.B[0] = 1. / 3.; rk.B[1] = 15. / 16.; rk.B[2] = 8. / 15.;
      rk.C[0] = 0.; rk.C[1] = 1. / 3.; rk.C[2] = 3. / 4.; rk.C[3] = 1.;
      break;
    case 4:
      rk.nstages = 5;
      rk.A[0] = 0.;
      rk.A[1] = -567301805773. / 1357537059087.;
      rk.A[2] = -2404267990393. / 2016746695238.;
      rk.A[3] = -3550918686646. / 2091501179385.;
      rk.A[4] = -1275806237668. / 842570457699.;
      rk.B[0] = 1432997174477. / 9575080441755.;
      rk.B[1] = 5161836677717. / 13612068292357.;
      rk.B[2] = 1720146321549. / 2090206949498.;
      rk.B[3] = 3134564353537. / 4481467310338.;
      rk.B[4] = 2277821191437. / 14882151754819.;
      rk.C[0] = 0.;
      rk.C[1] = 1432997174477. / 9575080441755.;
      rk.C[2] = 2526269341429. / 6820363962896.;
      rk.C[3] = 2006345519317. / 3224310063776.;
      rk.C[4] = 2802321613138. / 2924317926251.;
      rk.C[5] = 1.;
      break;
    default:
      throw Err("Error in initializing RK method");
  }
}

// ---------------------------------------------------------------------------
// perturbation shapes, pert.cpp:7-148
struct Pert {
  int type = 0;
  double t0 = 0., x0 = 0., dx = 0., y0 = 0., dy = 0.;
  double a = 0., b = 0., c = 0., d = 0.;
  double p[9] = {};  // law-specific amplitudes (load.cpp, swparam.cpp, stzparam.cpp)

  void init(const std::string& t, double t0i, double x0i, double dxi, double y0i, double dyi,
            const int n[2], const int xm[2], const int xm_loc[2], const double x[2], const double l[2]) {
    if (t == "constant") type = 0;
    else if (t == "gaussian") type = 1;
    else if (t == "ellipse") type = 2;
    else if (t == "linear") type = 3;
    else if (t == "boxcar") type = 4;
    else throw Err("unknown perturbation type " + t);
    if (t0i < 0. || dxi < 0. || dyi < 0.) throw Err("invalid perturbation parameters");
    t0 = t0i; x0 = x0i; y0 = y0i; dx = dxi; dy = dyi;
    a = (double)(xm_loc[0] - xm[0]) / (double)(n[0] - 1) * l[0] + x[0];
    b = l[0] / (double)(n[0] - 1);
    c = (double)(xm_loc[1] - xm[1]) / (double)(n[1] - 1) * l[1] + x[1];
    d = l[1] / (double)(n[1] - 1);
  }
  double tfunc(double t) const { return (t >= t0) ? 1. : t / t0; }
  double xc(int i) const { return b * (double)i + a; }
  double yc(int j) const { return d * (double)j + c; }
  double xyfunc(int i, int j) const {
    const double eps = 1.e-14;
    double xval, yval, xy = 0.;
    switch (type) {
      case 0:
        xy = 1.;
        break;
      case 1: {
        if (dx < eps) xval = 0.; else { double q = (xc(i) - x0) / dx; xval = q * q; }
        if (dy < eps) yval = 0.; else { double q = (yc(j) - y0) / dy; yval = q * q; }
        xy = exp(-xval - yval);
        break;
      }
      case 2: {
        if (dx < eps) xval = 0.; else { double q = (xc(i) - x0) / dx; xval = q * q; }
        if (dy < eps) yval = 0.; else { double q = (yc(j) - y0) / dy; yval = q * q; }
        if (xval + yval <= 1.) xy = 1.;
        break;
      }
      case 3:
        xval = (dx < eps) ? 1. : xc(i) / dx + x0;
        yval = (dy < eps) ? 1. : yc(j) / dy + y0;
        xy = xval * yval;
        break;
      case 4:
        if (dx < eps) xval = 1.;
        else if (fabs(fabs(xc(i) - x0) - dx) < eps) xval = 0.5;
        else if (fabs(xc(i) - x0) < dx) xval = 1.;
        else xval = 0.;
        if (dy < eps) yval = 1.;
        else if (fabs(fabs(yc(j) - y0) - dy) < eps) yval = 0.5;
        else if (fabs(yc(j) - y0) < dy) yval = 1.;
        else yval = 0.;
        xy = xval * yval;
        break;
    }
    return xy;
  }
  // amplitude k at (i,j,t): p[k]*xyfunc*tfunc (load.cpp:20-37 and friends)
  double get(int k, int i, int j, double t) const { return p[k] * xyfunc(i, j) * tfunc(t); }
};

// ---------------------------------------------------------------------------
// rotations, boundary.cpp:505-568
struct BF {
  double v1, v2, v3, s11, s12, s13, s22, s23, s33;
};

static BF rotate_xy_nt(const BF& b, const double nn[3], const double t1[3], const double t2[3]) {
  BF o;
  o.v1 = nn[0] * b.v1 + nn[1] * b.v2 + nn[2] * b.v3;
  o.v2 = t1[0] * b.v1 + t1[1] * b.v2 + t1[2] * b.v3;
  o.v3 = t2[0] * b.v1 + t2[1] * b.v2 + t2[2] * b.v3;
  o.s11 = (nn[0] * (b.s11 * nn[0] + b.s12 * nn[1] + b.s13 * nn[2]) +
           nn[1] * (b.s12 * nn[0] + b.s22 * nn[1] + b.s23 * nn[2]) +
           nn[2] * (b.s13 * nn[0] + b.s23 * nn[1] + b.s33 * nn[2]));
  o.s12 = (nn[0] * (b.s11 * t1[0] + b.s12 * t1[1] + b.s13 * t1[2]) +
           nn[1] * (b.s12 * t1[0] + b.s22 * t1[1] + b.s23 * t1[2]) +
           nn[2] * (b.s13 * t1[0] + b.s23 * t1[1] + b.s33 * t1[2]));
  o.s13 = (nn[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) +
           nn[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           nn[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  o.s22 = (t1[0] * (b.s11 * t1[0] + b.s12 * t1[1] + b.s13 * t1[2]) +
           t1[1] * (b.s12 * t1[0] + b.s22 * t1[1] + b.s23 * t1[2]) +
           t1[2] * (b.s13 * t1[0] + b.s23 * t1[1] + b.s33 * t1[2]));
  o.s23 = (t1[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) +
           t1[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           t1[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  o.s33 = (t2[0] * (b.s11 * t2[0] + b.s12 * t2[1] + b.s13 * t2[2]) +
           t2[1] * (b.s12 * t2[0] + b.s22 * t2[1] + b.s23 * t2[2]) +
           t2[2] * (b.s13 * t2[0] + b.s23 * t2[1] + b.s33 * t2[2]));
  return o;
}

static BF rotate_nt_xy(const BF& b, const double nn[3], const double t1[3], const double t2[3]) {
  BF o;
  o.v1 = nn[0] * b.v1 + t1[0] * b.v2 + t2[0] * b.v3;
  o.v2 = nn[1] * b.v1 + t1[1] * b.v2 + t2[1] * b.v3;
  o.v3 = nn[2] * b.v1 + t1[2] * b.v2 + t2[2] * b.v3;
  o.s11 = (nn[0] * (b.s11 * nn[0] + b.s12 * t1[0] + b.s13 * t2[0]) +
           t1[0] * (b.s12 * nn[0] + b.s22 * t1[0] + b.s23 * t2[0]) +
           t2[0] * (b.s13 * nn[0] + b.s23 * t1[0] + b.s33 * t2[0]));
  o.s12 = (nn[0] * (b.s11 * nn[1] + b.s12 * t1[1] + b.s13 * t2[1]) +
           t1[0] * (b.s12 * nn[1] + b.s22 * t1[1] + b.s23 * t2[1]) +
           t2[0] * (b.s13 * nn[1] + b.s23 * t1[1] + b.s33 * t2[1]));
  o.s13 = (nn[0] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) +
           t1[0] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[0] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  o.s22 = (nn[1] * (b.s11 * nn[1] + b.s12 * t1[1] + b.s13 * t2[1]) +
           t1[1] * (b.s12 * nn[1] + b.s22 * t1[1] + b.s23 * t2[1]) +
           t2[1] * (b.s13 * nn[1] + b.s23 * t1[1] + b.s33 * t2[1]));
  o.s23 = (nn[1] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) +
           t1[1] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[1] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  o.s33 = (nn[2] * (b.s11 * nn[2] + b.s12 * t1[2] + b.s13 * t2[2]) +
           t1[2] * (b.s12 * nn[2] + b.s22 * t1[2] + b.s23 * t2[2]) +
           t2[2] * (b.s13 * nn[2] + b.s23 * t1[2] + b.s33 * t2[2]));
  return o;
}

// ---------------------------------------------------------------------------
// STZ friction pieces: stz.cpp:436-484 and the bracketed Newton of
// utilities.cpp:89-180 (Numerical Recipes rtsafe with an overflow repair)
static double calc_vpl(double mu, const double* p) {
  double chi = p[3], v0 = p[4], f0 = p[5], a = p[6], muy = p[7];
  if (!(chi > 0.)) throw Err("STZ effective temperature must be positive");
  if (mu <= muy) return 0.;
  return v0 * exp(-f0 + mu / a - 1. / chi) * (1. - muy / mu);
}
static double calc_dvpldmu(double mu, const double* p) {
  double a = p[6], muy = p[7];
  if (mu <= muy) return 0.;
  return calc_vpl(mu, p) * (1. / a - muy / mu / (mu - muy));
}
static double stz_func(double mu, const double* p) { return p[1] * calc_vpl(mu, p) - p[2] * mu - p[0]; }
static double stz_der(double mu, const double* p) { return p[1] * calc_dvpldmu(mu, p) - p[2]; }

static double solve_newton(double xmin, double xmax, const double* params) {
  int i = 0;
  const int nmax = 100;
  const double tol = 2.e-16;
  double func, der, x, xl, xh, fh, dx, dxold, temp;
  x = 0.5 * (xmin + xmax);
  xh = xmax;
  xl = xmin;
  func = stz_func(x, params);
  fh = stz_func(xh, params);
  if (std::isinf(fh)) {
    while (true) {
      if (!std::isinf(func)) {
        if (func > 0.) { xh = x; break; }
        else xl = x;
      } else {
        xh = x;
      }
      x = 0.5 * (xh + xl);
      func = stz_func(x, params);
    }
  }
  x = 0.5 * (xh + xl);
  func = stz_func(x, params);
  der = stz_der(x, params);
  dx = fabs(xh - xl);
  dxold = dx;
  while (i < nmax) {
    if (der == 0.) throw Err("zero derivative in Newton's method in utilities.cpp");
    if ((((x - xh) * der - func) * ((x - xl) * der - func) > 0.) || fabs(2. * func) > fabs(dxold * der)) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      x = xl + dx;
      if (xl == x) break;
    } else {
      dxold = dx;
      dx = func / der;
      temp = x;
      x = x - dx;
      if (temp == x) break;
    }
    if (fabs(dx) < tol) break;
    func = stz_func(x, params);
    der = stz_der(x, params);
    if (func < 0.) xl = x; else xh = x;
    i++;
  }
  if (i == nmax) throw Err("Newton's method failed to converge in utilities.cpp");
  return x;
}

// ---------------------------------------------------------------------------
// plasticity, block.cpp:591-697
struct PS {
  double sxx, sxy, sxz, syy, syz, szz, gammap, lambda, epxx, epxy, epxz, epyy, epyz, epzz;
};
struct PMat {
  double mu, c, beta, eta;
};

static PS plastic_flow(double dt, const PS& in, double k, double g, const PMat& m) {
  double sigma = (in.sxx + in.syy + in.szz) / 3.;
  double tau = sqrt(((in.sxx - in.syy) * (in.sxx - in.syy) + (in.syy - in.szz) * (in.syy - in.szz) +
                     (in.szz - in.sxx) * (in.szz - in.sxx)) / 6. +
                    in.sxy * in.sxy + in.sxz * in.sxz + in.syz * in.syz);
  double yf = m.c - m.mu * sigma;
  if (yf < 0.) yf = 0.;
  double y = tau - yf;
  PS out;
  if (y <= 0.) {
    out = in;
    out.lambda = 0.;
    return out;
  }
  out = in;  // the reference leaves unused plastic-strain members undefined; copy keeps them defined
  if (tau * m.mu * m.beta * k > (m.mu * sigma - m.c) * g) {
    out.lambda = (tau + m.mu * sigma - m.c) / (m.eta + dt * (g + m.mu * m.beta * k));
  } else {
    out.lambda = tau / (m.eta + dt * g);
  }
  out.gammap = in.gammap + dt * out.lambda;
  double sd[6] = {in.sxx - sigma, in.sxy, in.sxz, in.syy - sigma, in.syz, in.szz - sigma};
  tau -= dt * out.lambda * g;
  sigma -= dt * out.lambda * m.beta * k;
  for (int i = 0; i < 6; i++) sd[i] *= tau / (tau + dt * out.lambda * g);
  out.sxx = sd[0] + sigma;
  out.sxy = sd[1];
  out.sxz = sd[2];
  out.syy = sd[3] + sigma;
  out.syz = sd[4];
  out.szz = sd[5] + sigma;
  out.epxx = in.epxx + dt * out.lambda * (sd[0] / (2. * tau) + (m.beta / 3.));
  out.epxy = in.epxy + dt * out.lambda * (sd[1] / (2. * tau));
  out.epxz = in.epxz + dt * out.lambda * (sd[2] / (2. * tau));
  out.epyy = in.epyy + dt * out.lambda * (sd[3] / (2. * tau) + (m.beta / 3.));
  out.epyz = in.epyz + dt * out.lambda * (sd[4] / (2. * tau));
  out.epzz = in.epzz + dt * out.lambda * (sd[5] / (2. * tau) + (m.beta / 3.));
  return out;
}

// ---------------------------------------------------------------------------
// problem state

struct Bound {  // boundary.cpp:15-240
  bool active = false;
  int loc = 0;
  double r = 0., alpha = 0., theta = 0.;
  double cp = 0., cs = 0., zp = 0., zs = 0., gamma = 0.;
  int n0 = 0, n1 = 0;
  int mlb[3] = {0, 0, 0}, prb[3] = {0, 0, 0};
  std::vector<double> nrm;  // [ndim][n0][n1]
  std::vector<double> dl;   // [n0][n1]
};

struct Block {
  int coords[3] = {0, 0, 0};
  int nx[3] = {1, 1, 1}, xm[3] = {0, 0, 0};
  double rho = 0., lam = 0., G = 0.;
  PMat pm{0.5735, 0., 0.2867, 0.2775};  // material.cpp:15-21 defaults
  double xb[3] = {0., 0., 0.}, lb[3] = {0., 0., 0.};
  std::string btype[6], bfile[6];
  int mlb[3], mc[3], mrb[3], prb[3];
  double dx[3];
  double cp, cs, zp, zs;
  Bound bnd[6];
};

struct Iface {
  int id = 0, type = 0;  // 0 locked, 1 frictionless, 2 slipweak, 3 stz
  int dir = 0;
  int b1[3], b2[3];
  int n[2];
  int mlb[3], prb[3], delta[3];
  int xm[3], xp[3], xm_loc[3], xp_loc[3];
  double x[3] = {0., 0., 0.}, l[3] = {0., 0., 0.};
  double cp1, cs1, zp1, zs1, cp2, cs2, zp2, zs2, gamma1, gamma2;
  std::vector<double> nrm, dl1, dl2;
  bool is_friction = false, has_state = false;
  // friction
  std::vector<Pert> loads;
  bool load_file = false;
  std::vector<double> ls1, ls2, ls3;
  std::vector<Pert> perts;
  bool param_file = false;
  std::vector<double> prm[9];
  double chi0 = 0.;
  std::vector<double> u, du, v, s, sn, ux, dux, vx, sx, state, dstate, dstatedt;
};

struct OutUnit {  // outputunit.cpp:13-741 (single rank)
  std::string name, field_name;
  int field = 0, location = -1;
  int tm = 0, tp = 0, ts = 1;
  int xm[3], xp[3], xs[3], nx[3];
  int comp = 0;  // component offset for interface vector arrays
  FILE* fp = nullptr;
  FILE* tfp = nullptr;
};

struct Front {  // front.cpp:15-125
  int iface = 0;
  int field = 0;  // 0 V, 1 U (never triggers, front.cpp:116-120)
  double value = 0.;
  std::vector<double> tvals;
};

struct Problem {
  std::string inpath;
  std::string name, datadir;
  int nt = 0;
  double dt = 0., ttot = 0., cfl = 0.;
  int ninfo = 1, rkorder = 0;
  RK rk;
  FD fd;
  int ndim = 0, mode = 0, nx[3] = {1, 1, 1}, nblocks[3] = {1, 1, 1};
  std::vector<int> nxb[3], xmb[3];
  int nifaces = 0;
  std::vector<std::string> iftype;
  int sbporder = 0;
  bool plastic = false, plastic_tensor = false;
  double s0[6] = {0., 0., 0., 0., 0., 0.};
  std::string loadfile, matfile;
  bool hetstress = false, hetmat = false;
  int nf = 0, nfp = 0, ns = 0, nv = 0, nmat = 0, sidx[6] = {0, 1, 2, 3, 4, 5};
  size_t N = 0;
  int Ny = 1, Nz = 1;  // strides: index = i*Ny*Nz + j*Nz + k
  std::vector<double> f, df, x, metric, jac, s, mat;
  double cdiss = 0.;
  bool dissipation = false;
  std::vector<Block> blocks;
  std::vector<Iface> ifaces;
  std::vector<OutUnit> outs;
  std::vector<Front> fronts;
  int step = 0;

  size_t idx(int i, int j, int k) const { return (size_t)i * Ny * Nz + (size_t)j * Nz + k; }
  Block& blk(int i, int j, int k) { return blocks[(i * nblocks[1] + j) * nblocks[2] + k]; }
};

// ---------------------------------------------------------------------------
// setup

static void read_problem(Problem& P, InText& in) {  // problem.cpp:16-58
  in.rewind();
  if (!in.scan("[fdfault.problem]")) throw Err("Error reading problem from input file");
  P.name = in.tok();
  P.datadir = in.tok();
  P.nt = in.integer();
  P.dt = in.dbl();
  P.ttot = in.dbl();
  P.cfl = in.dbl();
  P.ninfo = in.integer();
  P.rkorder = in.integer();
  rk_init(P.rk, P.rkorder);
}

static void read_domain(Problem& P, InText& in) {  // domain.cpp:19-153
  in.rewind();
  if (!in.scan("[fdfault.domain]")) throw Err("Error reading domain from input file");
  P.ndim = in.integer();
  P.mode = in.integer();
  for (int i = 0; i < 3; i++) P.nx[i] = in.integer();
  for (int i = 0; i < 3; i++) P.nblocks[i] = in.integer();
  for (int i = 0; i < 3; i++) {
    if (P.nblocks[i] <= 0) throw Err("invalid number of blocks");
    P.nxb[i].resize(P.nblocks[i]);
    P.xmb[i].resize(P.nblocks[i]);
    for (int j = 0; j < P.nblocks[i]; j++) P.nxb[i][j] = in.integer();
  }
  P.nifaces = in.integer();
  for (int i = 0; i < P.nifaces; i++) P.iftype.push_back(in.tok());
  P.sbporder = in.integer();
  std::string material = in.tok();
  if (P.ndim != 2 && P.ndim != 3) throw Err("ndim must be 2 or 3");
  if (P.mode != 2 && P.mode != 3) throw Err("mode must be 2 or 3");
  if (material != "elastic" && material != "plastic") throw Err("material must be elastic or plastic");
  if (P.ndim == 2 && P.nx[2] != 1) throw Err("2d problems require nz = 1");
  P.plastic = (material == "plastic");
  for (int i = 0; i < 3; i++) {
    int sum = 0;
    for (int j = 0; j < P.nblocks[i]; j++) {
      sum += P.nxb[i][j];
      P.xmb[i][j] = (j == 0) ? 0 : P.xmb[i][j - 1] + P.nxb[i][j - 1];
    }
    if (sum != P.nx[i]) throw Err("block sizes do not sum to grid size");
  }
  fd_init(P.fd, P.sbporder);
}

static void read_fields(Problem& P, InText& in) {  // fields.cpp:13-164
  in.rewind();
  if (!in.scan("[fdfault.fields]")) throw Err("Error reading fields from input file");
  for (int i = 0; i < 6; i++) P.s0[i] = in.dbl();
  P.loadfile = in.tok();
  P.matfile = in.tok();
  P.plastic_tensor = in.integer() != 0;
  if (P.ndim == 3) { P.nf = 9; P.nmat = 3; }
  else if (P.mode == 2) { P.nf = 5; P.nmat = 3; }
  else { P.nf = 3; P.nmat = 2; }
  if (!P.plastic) P.nfp = 0;
  else {
    P.nfp = P.plastic_tensor ? 8 : 2;
    if (P.ndim == 2 && P.mode == 2) P.nf = 6;
  }
  if (P.ndim == 3) {
    P.ns = 6; P.nv = 3;
    for (int i = 0; i < 6; i++) P.sidx[i] = i;
  } else if (P.mode == 2) {
    P.ns = 3; P.sidx[0] = 0; P.sidx[1] = 1; P.sidx[2] = 3;
    if (P.plastic) { P.ns = 4; P.sidx[3] = 5; }
    P.nv = 2;
  } else {
    P.ns = 2; P.nv = 1; P.sidx[0] = 2; P.sidx[1] = 4;
  }
  P.Ny = P.nx[1];
  P.Nz = P.nx[2];
  P.N = (size_t)P.nx[0] * P.nx[1] * P.nx[2];
  P.f.assign((size_t)(P.nf + P.nfp) * P.N, 0.);
  P.df.assign((size_t)P.nf * P.N, 0.);
  P.x.assign((size_t)P.ndim * P.N, 0.);
  P.metric.assign((size_t)P.ndim * P.ndim * P.N, 0.);
  P.jac.assign(P.N, 0.);
  if (P.loadfile != "none") {
    P.hetstress = true;
    P.s = read_binary(P.loadfile, (size_t)P.ns * P.N, "Error opening stress file in fields.cpp",
                      "heterogeneous stress");
  }
  if (P.matfile != "none") {
    P.hetmat = true;
    P.mat = read_binary(P.matfile, (size_t)P.nmat * P.N, "Error opening material file in fields.cpp",
                        "heterogeneous material");
  }
}

// flat or file surface, surface.cpp:11-101 ; returns [ndim][n0][n1]
static std::vector<double> make_surface(const Problem& P, const Block& B, int face, int& n0, int& n1) {
  int dir = face / 2;
  int id0 = (dir == 0) ? 1 : 0;
  int id1 = (dir == 2) ? 1 : 2;
  n0 = B.nx[id0];
  n1 = B.nx[id1];
  size_t n = (size_t)n0 * n1;
  if (B.bfile[face] != "none")
    return read_binary(B.bfile[face], (size_t)P.ndim * n, "Error reading surface from file " + B.bfile[face],
                       "surface");
  double xin[3], lin[2];
  for (int i = 0; i < 3; i++) xin[i] = B.xb[i];
  if (face % 2 == 1) xin[dir] = B.xb[dir] + B.lb[dir];
  static const int lmap[6][2] = {{1, 2}, {1, 2}, {0, 2}, {0, 2}, {0, 1}, {0, 1}};
  lin[0] = B.lb[lmap[face][0]];
  lin[1] = B.lb[lmap[face][1]];
  if (lin[0] < 0. || lin[1] < 0.) throw Err("negative block length");
  std::vector<double> xs((size_t)P.ndim * n, 0.);
  for (int j = 0; j < n0; j++)
    for (int k = 0; k < n1; k++) xs[dir * n + j * n1 + k] = xin[dir];
  for (int j = 0; j < n0; j++)
    for (int k = 0; k < n1; k++) {
      xs[id0 * n + j * n1 + k] = xin[id0] + lin[0] * (double)(j) / (double)(n0 - 1);
      if (P.ndim == 3) xs[id1 * n + j * n1 + k] = xin[id1] + lin[1] * (double)(k) / (double)(n1 - 1);
    }
  return xs;
}

struct Surf {
  std::vector<double> x;
  int n0, n1;
  double get(int i, int a, int b) const { return x[(size_t)i * n0 * n1 + (size_t)a * n1 + b]; }
};

// surface::has_same_edge, surface.cpp:129-247 (tolerance check of shared block edges)
static bool same_edge(const Problem& P, const Surf& s1, int e1, const Surf& s2, int e2) {
  const double abtol = 1.e-8, reltol = 1.e-5;
  // edge 0: second index 0; 1: first index 0; 2: second index n1-1; 3: first index n0-1
  auto edge = [&](const Surf& s, int e, int i, int m) -> double {
    switch (e) {
      case 0: return s.get(i, m, 0);
      case 1: return s.get(i, 0, m);
      case 2: return s.get(i, m, s.n1 - 1);
      default: return s.get(i, s.n0 - 1, m);
    }
  };
  int len1 = (e1 % 2 == 1) ? s1.n1 : s1.n0;
  int len2 = (e2 % 2 == 1) ? s2.n1 : s2.n0;
  if (len1 != len2) return false;
  for (int i = 0; i < P.ndim; i++)
    for (int m = 0; m < len1; m++) {
      double a = edge(s1, e1, i, m), b = edge(s2, e2, i, m);
      if (fabs(a - b) > abtol + reltol * fabs(b)) return false;
    }
  return true;
}

// block::set_grid, block.cpp:769-1046
static void set_grid(Problem& P, Block& B, const Surf* sf) {
  for (int i = 0; i < 3; i++) B.dx[i] = 1. / (double)(B.nx[i] - 1);
  const int nx = B.nx[0], ny = B.nx[1], nz = B.nx[2];
  for (int i = 0; i < P.ndim; i++)
    for (int j = B.mlb[0]; j < B.prb[0]; j++)
      for (int k = B.mlb[1]; k < B.prb[1]; k++)
        for (int l = B.mlb[2]; l < B.prb[2]; l++) {
          int jj = j - B.mlb[0], kk = k - B.mlb[1], ll = l - B.mlb[2];
          double p = (double)jj * B.dx[0], q = (double)kk * B.dx[1], r = (double)ll * B.dx[2];
          double& X = P.x[(size_t)i * P.N + P.idx(j, k, l)];
          X = ((1. - p) * sf[0].get(i, kk, ll) + p * sf[1].get(i, kk, ll) + (1. - q) * sf[2].get(i, jj, ll) +
               q * sf[3].get(i, jj, ll));
          if (P.ndim == 3) X += (1. - r) * sf[4].get(i, jj, kk) + r * sf[5].get(i, jj, kk);
          X -= ((1. - q) * (1. - p) * sf[0].get(i, 0, ll) + (1. - q) * p * sf[1].get(i, 0, ll) +
                q * (1. - p) * sf[0].get(i, ny - 1, ll) + q * p * sf[1].get(i, ny - 1, ll));
          if (P.ndim == 3) {
            X -= ((1. - p) * (1. - r) * sf[0].get(i, kk, 0) + p * (1. - r) * sf[1].get(i, kk, 0) +
                  (1. - q) * (1. - r) * sf[2].get(i, jj, 0) + q * (1. - r) * sf[3].get(i, jj, 0) +
                  (1. - p) * r * sf[0].get(i, kk, nz - 1) + p * r * sf[1].get(i, kk, nz - 1) +
                  (1. - q) * r * sf[2].get(i, jj, nz - 1) + q * r * sf[3].get(i, jj, nz - 1));
            X += ((1. - p) * (1. - q) * (1. - r) * sf[0].get(i, 0, 0) + p * (1. - q) * (1. - r) * sf[1].get(i, 0, 0) +
                  (1. - p) * q * (1. - r) * sf[0].get(i, ny - 1, 0) + (1. - p) * (1. - q) * r * sf[0].get(i, 0, nz - 1) +
                  p * q * (1. - r) * sf[1].get(i, ny - 1, 0) + p * (1. - q) * r * sf[1].get(i, 0, nz - 1) +
                  (1. - p) * q * r * sf[0].get(i, ny - 1, nz - 1) + p * q * r * sf[1].get(i, ny - 1, nz - 1));
          }
        }
  (void)nx;
  // coordinate derivatives xp[l][d] at each block point (identity where unused)
  const size_t nb = (size_t)B.nx[0] * B.nx[1] * B.nx[2];
  std::vector<double> xp(9 * nb, 0.);
  auto XP = [&](int l, int d, int i, int j, int k) -> double& {
    return xp[(size_t)(l * 3 + d) * nb + ((size_t)(i - B.mlb[0]) * B.nx[1] + (j - B.mlb[1])) * B.nx[2] + (k - B.mlb[2])];
  };
  for (int l = 0; l < 3; l++)
    for (size_t q = 0; q < nb; q++) xp[(size_t)(l * 3 + l) * nb + q] = 1.;
  const int s = P.fd.s;
  for (int d = 0; d < P.ndim; d++) {
    for (int i = B.mlb[0]; i < B.prb[0]; i++)
      for (int j = B.mlb[1]; j < B.prb[1]; j++)
        for (int k = B.mlb[2]; k < B.prb[2]; k++) {
          int c[3] = {i, j, k};
          int id = c[d];
          for (int l = 0; l < P.ndim; l++) {
            double acc = 0.;
            auto xat = [&](int m) {
              int cc[3] = {i, j, k};
              cc[d] = m;
              return P.x[(size_t)l * P.N + P.idx(cc[0], cc[1], cc[2])];
            };
            if (id < B.mc[d]) {
              for (int n = 0; n < 3 * (s - 1); n++) acc += P.fd.c[id - B.mlb[d] + 1][n] * xat(B.mlb[d] + n) / B.dx[d];
            } else if (id < B.mrb[d]) {
              for (int n = 0; n < 2 * s - 1; n++) acc += P.fd.c[0][n] * xat(id - s + 1 + n) / B.dx[d];
            } else {
              for (int n = 0; n < 3 * (s - 1); n++) acc -= P.fd.c[B.prb[d] - id][n] * xat(B.prb[d] - 1 - n) / B.dx[d];
            }
            XP(l, d, i, j, k) = acc;
          }
        }
  }
  for (int i = B.mlb[0]; i < B.prb[0]; i++)
    for (int j = B.mlb[1]; j < B.prb[1]; j++)
      for (int k = B.mlb[2]; k < B.prb[2]; k++) {
        size_t id = P.idx(i, j, k);
        auto X = [&](int a, int b) { return XP(a, b, i, j, k); };
        double J = (X(0, 0) * (X(1, 1) * X(2, 2) - X(1, 2) * X(2, 1)) - X(1, 0) * (X(0, 1) * X(2, 2) - X(0, 2) * X(2, 1)) +
                    X(2, 0) * (X(0, 1) * X(1, 2) - X(0, 2) * X(1, 1)));
        if (!(J > 0.)) throw Err("non-positive Jacobian in block grid (grid is not smooth enough)");
        P.jac[id] = J;
        for (int l = 0; l < P.ndim; l++)
          for (int m = 0; m < P.ndim; m++)
            P.metric[(size_t)(l * P.ndim + m) * P.N + id] =
                ((X((m + 1) % 3, (l + 1) % 3) * X((m + 2) % 3, (l + 2) % 3) - X((m + 1) % 3, (l + 2) % 3) * X((m + 2) % 3, (l + 1) % 3)) / J);
      }
}

// boundary::allocate_normals, boundary.cpp:173-240 (outward normal, h scale)
static void boundary_normals(Problem& P, Block& B, Bound& bd) {
  int loc = bd.loc, dir = loc / 2;
  size_t n = (size_t)bd.n0 * bd.n1;
  bd.nrm.assign((size_t)P.ndim * n, 0.);
  bd.dl.assign(n, 0.);
  for (int i = 0; i < bd.n0; i++)
    for (int j = 0; j < bd.n1; j++) {
      size_t id;
      if (dir == 0) id = P.idx(bd.mlb[0], i + bd.mlb[1], j + bd.mlb[2]);
      else if (dir == 1) id = P.idx(i + bd.mlb[0], bd.mlb[1], j + bd.mlb[2]);
      else id = P.idx(i + bd.mlb[0], j + bd.mlb[1], bd.mlb[2]);
      double dl = 0.;
      for (int k = 0; k < P.ndim; k++) {
        double m = P.metric[(size_t)(dir * P.ndim + k) * P.N + id];
        dl += m * m;
        bd.nrm[k * n + i * bd.n1 + j] = m;
      }
      dl = sqrt(dl);
      for (int k = 0; k < P.ndim; k++) bd.nrm[k * n + i * bd.n1 + j] /= dl;
      dl /= P.fd.h0 * B.dx[dir];
      bd.dl[i * bd.n1 + j] = dl;
      if (loc % 2 == 0)
        for (int k = 0; k < P.ndim; k++) bd.nrm[k * n + i * bd.n1 + j] = -bd.nrm[k * n + i * bd.n1 + j];
    }
}

static void read_blocks(Problem& P, InText& in) {  // block.cpp:22-264, domain.cpp:322-358
  in.rewind();
  P.cdiss = 0.;
  P.dissipation = false;
  if (in.scan("[fdfault.operator]")) {
    P.cdiss = in.dbl();
    P.dissipation = P.cdiss > 0.;
  }
  P.blocks.resize((size_t)P.nblocks[0] * P.nblocks[1] * P.nblocks[2]);
  const int s = P.sbporder;
  for (int bi = 0; bi < P.nblocks[0]; bi++)
    for (int bj = 0; bj < P.nblocks[1]; bj++)
      for (int bk = 0; bk < P.nblocks[2]; bk++) {
        Block& B = P.blk(bi, bj, bk);
        B.coords[0] = bi; B.coords[1] = bj; B.coords[2] = bk;
        B.nx[0] = P.nxb[0][bi]; B.nx[1] = P.nxb[1][bj]; B.nx[2] = P.nxb[2][bk];
        B.xm[0] = P.xmb[0][bi]; B.xm[1] = P.xmb[1][bj]; B.xm[2] = P.xmb[2][bk];
        std::string hdr = "[fdfault.block" + std::to_string(bi) + std::to_string(bj) + std::to_string(bk) + "]";
        in.rewind();
        if (!in.scan(hdr)) throw Err("Error reading block " + hdr.substr(14, hdr.size() - 15) + " from input file");
        B.rho = in.dbl();
        B.lam = in.dbl();
        B.G = in.dbl();
        if (P.plastic) {
          B.pm.mu = in.dbl();
          B.pm.c = in.dbl();
          B.pm.beta = in.dbl();
          B.pm.eta = in.dbl();
        }
        for (int i = 0; i < P.ndim; i++) B.xb[i] = in.dbl();
        for (int i = 0; i < P.ndim; i++) B.lb[i] = in.dbl();
        for (int i = 0; i < 2 * P.ndim; i++) B.btype[i] = in.tok();
        for (int i = 0; i < 2 * P.ndim; i++) B.bfile[i] = in.tok();
        for (int i = 0; i < P.ndim; i++)
          if (B.nx[i] < 4 * s - 3) throw Err("block too small for the finite difference method");
        // single-rank index ranges, block.cpp:165-179
        for (int i = 0; i < 3; i++) {
          B.mlb[i] = B.xm[i];
          B.mc[i] = (B.nx[i] > 1) ? B.mlb[i] + 2 * (s - 1) : B.mlb[i];
          if (B.nx[i] > 1) {
            B.mrb[i] = B.mlb[i] + B.nx[i] - 2 * (s - 1);
            B.prb[i] = B.mrb[i] + 2 * (s - 1);
          } else {
            B.mrb[i] = B.mlb[i] + B.nx[i];
            B.prb[i] = B.mrb[i];
          }
        }
        B.cs = sqrt(B.G / B.rho);
        B.cp = sqrt((B.lam + 2. * B.G) / B.rho);
        B.zs = B.rho * B.cs;
        B.zp = B.rho * B.cp;
        Surf sf[6];
        int nb = 2 * P.ndim;
        for (int i = 0; i < nb; i++) sf[i].x = make_surface(P, B, i, sf[i].n0, sf[i].n1);
        static const int surf1[12] = {0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 3, 3};
        static const int surf2[12] = {2, 3, 2, 3, 4, 5, 4, 5, 4, 5, 4, 5};
        static const int edge1[12] = {1, 3, 1, 3, 0, 2, 0, 2, 0, 2, 0, 2};
        static const int edge2[12] = {1, 1, 3, 3, 1, 1, 3, 3, 0, 0, 2, 2};
        int nedge = (P.ndim == 3) ? 12 : 4;
        for (int e = 0; e < nedge; e++)
          if (!same_edge(P, sf[surf1[e]], edge1[e], sf[surf2[e]], edge2[e]))
            throw Err("Surface edges do not match in block.cpp");
        set_grid(P, B, sf);
      }
  // boundaries after all grids exist (each uses its own block's metric only)
  for (Block& B : P.blocks) {
    for (int loc = 0; loc < 2 * P.ndim; loc++) {
      Bound& bd = B.bnd[loc];
      bd.loc = loc;
      const std::string& bt = B.btype[loc];
      if (bt != "none" && bt != "absorbing" && bt != "free" && bt != "rigid") throw Err("unknown boundary type " + bt);
      if (bt == "none") continue;
      bd.active = true;
      int dir = loc / 2;
      for (int i = 0; i < 3; i++) bd.mlb[i] = B.mlb[i];
      int a = (dir == 0) ? 1 : 0, b = (dir == 2) ? 1 : 2;
      bd.n0 = B.nx[a];
      bd.n1 = B.nx[b];
      if (loc % 2 == 1) bd.mlb[dir] += B.nx[dir] - 1;
      bd.prb[dir] = bd.mlb[dir] + 1;
      bd.prb[a] = bd.mlb[a] + bd.n0;
      bd.prb[b] = bd.mlb[b] + bd.n1;
      bd.cp = B.cp; bd.cs = B.cs; bd.zp = B.zp; bd.zs = B.zs;
      double rr = bd.cs / bd.cp;
      bd.gamma = 1. - 2. * (rr * rr);
      if (bt == "absorbing") { bd.r = 0.; bd.alpha = 2.; bd.theta = 0.; }
      else if (bt == "free") { bd.r = -1.; bd.alpha = 1.; bd.theta = 0.; }
      else { bd.r = 1.; bd.alpha = 0.; bd.theta = 1.; }
      boundary_normals(P, B, bd);
    }
  }
}

// reads [3]/[6]/[9] x [n0][n1] parameter stacks (friction.cpp:507-567 and friends)
static void read_stack(const std::string& path, int nstack, size_t n, std::vector<double>* dst,
                       const std::string& open_msg, const char* what) {
  std::vector<double> all = read_binary(path, (size_t)nstack * n, open_msg, what);
  for (int k = 0; k < nstack; k++) dst[k].assign(all.begin() + k * n, all.begin() + (k + 1) * n);
}

static void read_interfaces(Problem& P, InText& in) {  // interface.cpp:17-403, friction/slipweak/stz ctors
  P.ifaces.resize(P.nifaces);
  for (int ii = 0; ii < P.nifaces; ii++) {
    Iface& I = P.ifaces[ii];
    I.id = ii;
    const std::string& t = P.iftype[ii];
    if (t == "locked") I.type = 0;
    else if (t == "frictionless") I.type = 1;
    else if (t == "slipweak") I.type = 2;
    else if (t == "stz") I.type = 3;
    else throw Err("unknown interface type " + t);
    std::string hdr = "[fdfault.interface" + std::to_string(ii) + "]";
    in.rewind();
    if (!in.scan(hdr)) throw Err("Error reading interface " + std::to_string(ii) + " from input file");
    std::string dirs = in.tok();
    for (int i = 0; i < 3; i++) I.b1[i] = in.integer();
    for (int i = 0; i < 3; i++) I.b2[i] = in.integer();
    if (dirs == "x") I.dir = 0;
    else if (dirs == "y") I.dir = 1;
    else if (dirs == "z" && P.ndim == 3) I.dir = 2;
    else throw Err("bad interface direction");
    for (int i = 0; i < 3; i++) {
      int want = I.b1[i] + (i == I.dir ? 1 : 0);
      if (I.b2[i] != want) throw Err("interface blocks are not adjacent");
      if (I.b1[i] < 0 || I.b2[i] >= P.nblocks[i]) throw Err("interface block index out of range");
    }
    Block& B1 = P.blk(I.b1[0], I.b1[1], I.b1[2]);
    Block& B2 = P.blk(I.b2[0], I.b2[1], I.b2[2]);
    const int d = I.dir;
    int a = (d == 0) ? 1 : 0, b = (d == 2) ? 1 : 2;
    if (B1.nx[a] != B2.nx[a] || B1.nx[b] != B2.nx[b]) throw Err("interface blocks do not match in size");
    I.n[0] = B1.nx[a];
    I.n[1] = B1.nx[b];
    for (int i = 0; i < 3; i++) {
      I.xm[i] = B2.xm[i];
      I.xp[i] = (i == d) ? B2.xm[i] : B2.xm[i] + B2.nx[i] - 1;
      I.xm_loc[i] = I.xm[i];
      I.xp_loc[i] = I.xp[i];
      I.mlb[i] = (i == d) ? B1.xm[i] + B1.nx[i] - 1 : B1.xm[i];
      I.delta[i] = (i == d) ? 1 : 0;
    }
    I.prb[d] = I.mlb[d] + 1;
    I.prb[a] = I.mlb[a] + I.n[0];
    I.prb[b] = I.mlb[b] + I.n[1];
    // matching edges, interface.cpp:333-347
    for (int i = 0; i < P.ndim; i++)
      for (int x0 = I.mlb[0]; x0 < I.prb[0]; x0++)
        for (int x1 = I.mlb[1]; x1 < I.prb[1]; x1++)
          for (int x2 = I.mlb[2]; x2 < I.prb[2]; x2++) {
            double va = P.x[(size_t)i * P.N + P.idx(x0, x1, x2)];
            double vb = P.x[(size_t)i * P.N + P.idx(x0 + I.delta[0], x1 + I.delta[1], x2 + I.delta[2])];
            if (!(fabs(va - vb) <= (1.e-8 + 1.e-5 * fabs(vb)))) throw Err("interface block edges do not match");
          }
    I.cp1 = B1.cp; I.cs1 = B1.cs; I.zp1 = B1.zp; I.zs1 = B1.zs;
    I.cp2 = B2.cp; I.cs2 = B2.cs; I.zp2 = B2.zp; I.zs2 = B2.zs;
    double r1 = I.cs1 / I.cp1, r2 = I.cs2 / I.cp2;
    I.gamma1 = 1. - 2. * (r1 * r1);
    I.gamma2 = 1. - 2. * (r2 * r2);
    for (int i = 0; i < 3; i++) { I.x[i] = B1.xb[i]; I.l[i] = B1.lb[i]; }
    // normals and grid spacings, interface.cpp:414-534
    size_t n = (size_t)I.n[0] * I.n[1];
    I.nrm.assign((size_t)P.ndim * n, 0.);
    I.dl1.assign(n, 0.);
    I.dl2.assign(n, 0.);
    for (int i = 0; i < I.n[0]; i++)
      for (int j = 0; j < I.n[1]; j++) {
        int c[3];
        c[d] = I.mlb[d];
        c[a] = i + I.mlb[a];
        c[b] = j + I.mlb[b];
        size_t id1 = P.idx(c[0], c[1], c[2]);
        c[d] += 1;
        size_t id2 = P.idx(c[0], c[1], c[2]);
        double dl = 0.;
        for (int k = 0; k < P.ndim; k++) {
          double m = P.metric[(size_t)(d * P.ndim + k) * P.N + id1];
          dl += m * m;
          I.nrm[k * n + i * I.n[1] + j] = m;
        }
        dl = sqrt(dl);
        for (int k = 0; k < P.ndim; k++) I.nrm[k * n + i * I.n[1] + j] /= dl;
        I.dl1[i * I.n[1] + j] = dl / (P.fd.h0 * B1.dx[d]);
        double dl2 = 0.;
        for (int k = 0; k < P.ndim; k++) {
          double m = P.metric[(size_t)(d * P.ndim + k) * P.N + id2];
          dl2 += m * m;
        }
        dl2 = sqrt(dl2);
        I.dl2[i * I.n[1] + j] = dl2 / (P.fd.h0 * B2.dx[d]);
      }
    if (I.type == 0) continue;
    // friction base, friction.cpp:19-227
    I.is_friction = true;
    I.ux.assign((size_t)(P.ndim - 1) * n, 0.);
    I.dux = I.ux; I.vx = I.ux; I.sx = I.ux;
    I.u.assign(n, 0.);
    I.du = I.u; I.v = I.u; I.s = I.u; I.sn = I.u;
    double x2d[2] = {I.x[a], I.x[b]}, l2d[2] = {I.l[a], I.l[b]};
    int xm2d[2] = {I.xm[a], I.xm[b]}, xml2d[2] = {I.xm_loc[a], I.xm_loc[b]};
    in.rewind();
    in.scan(hdr);
    if (!in.scan("[fdfault.friction]")) throw Err("Error reading interface " + std::to_string(ii) + " from input file");
    int nloads = in.integer();
    for (int k = 0; k < nloads; k++) {
      std::string lt = in.tok();
      double t0 = in.dbl(), x0 = in.dbl(), dx = in.dbl(), y0 = in.dbl(), dy = in.dbl();
      double s1 = in.dbl(), s2 = in.dbl(), s3 = in.dbl();
      if (P.ndim == 2 && P.mode == 2) s3 = 0.;
      else if (P.ndim == 2 && P.mode == 3) s2 = 0.;
      Pert pt;
      pt.init(lt, t0, x0, dx, y0, dy, I.n, xm2d, xml2d, x2d, l2d);
      pt.p[0] = s1; pt.p[1] = s2; pt.p[2] = s3;
      I.loads.push_back(pt);
    }
    std::string lf = in.tok();
    if (lf != "none") {
      I.load_file = true;
      std::vector<double> st[3];
      read_stack(lf, 3, n, st, "Error opening load file in friction.cpp", "load");
      I.ls1 = st[0]; I.ls2 = st[1]; I.ls3 = st[2];
      if (P.ndim == 2 && P.mode == 2) std::fill(I.ls3.begin(), I.ls3.end(), 0.);
      else if (P.ndim == 2 && P.mode == 3) std::fill(I.ls2.begin(), I.ls2.end(), 0.);
    }
    if (I.type == 2) {  // slipweak.cpp:17-205
      in.rewind();
      in.scan(hdr);
      if (!in.scan("[fdfault.slipweak]")) throw Err("Error reading slipweak from input file");
      int np = in.integer();
      for (int k = 0; k < np; k++) {
        std::string pt_t = in.tok();
        double t0 = in.dbl(), x0 = in.dbl(), dx = in.dbl(), y0 = in.dbl(), dy = in.dbl();
        Pert pt;
        pt.init(pt_t, t0, x0, dx, y0, dy, I.n, xm2d, xml2d, x2d, l2d);
        for (int q = 0; q < 6; q++) pt.p[q] = in.dbl();  // dc mus mud c0 trup tc
        I.perts.push_back(pt);
      }
      std::string pf = in.tok();
      if (pf != "none") { I.param_file = true; read_stack(pf, 6, n, I.prm, "Error opening file in friction.cpp", "slip weakening parameters"); }
    } else if (I.type == 3) {  // stz.cpp:1-230
      I.has_state = true;
      in.rewind();
      in.scan(hdr);
      if (!in.scan("[fdfault.stz]")) throw Err("Error reading stz from input file");
      I.chi0 = in.dbl();
      std::string sfile = in.tok();
      int np = in.integer();
      for (int k = 0; k < np; k++) {
        std::string pt_t = in.tok();
        double t0 = in.dbl(), x0 = in.dbl(), dx = in.dbl(), y0 = in.dbl(), dy = in.dbl();
        Pert pt;
        pt.init(pt_t, t0, x0, dx, y0, dy, I.n, xm2d, xml2d, x2d, l2d);
        for (int q = 0; q < 9; q++) pt.p[q] = in.dbl();  // v0 f0 a muy c0 R beta chiw v1
        I.perts.push_back(pt);
      }
      std::string pf = in.tok();
      I.state.assign(n, 0.);
      I.dstate.assign(n, 0.);
      I.dstatedt.assign(n, 0.);
      if (sfile != "none") I.state = read_binary(sfile, n, "Error opening state file in friction.cpp", "state");
      for (size_t q = 0; q < n; q++) I.state[q] += I.chi0;
      if (pf != "none") { I.param_file = true; read_stack(pf, 9, n, I.prm, "Error opening file in stz.cpp", "stz parameters"); }
    }
  }
}

// domain::get_min_dx / block::get_min_dx, domain.cpp:199-226, block.cpp:368-403
static double get_min_dx(const Problem& P) {
  double dxmin = 0.;
  for (const Block& B : P.blocks) {
    double dxmax = 0.;
    for (int i = B.mlb[0]; i < B.prb[0]; i++)
      for (int j = B.mlb[1]; j < B.prb[1]; j++)
        for (int k = B.mlb[2]; k < B.prb[2]; k++) {
          size_t id = P.idx(i, j, k);
          for (int l = 0; l < P.ndim; l++) {
            double t = 0.;
            for (int m = 0; m < P.ndim; m++) {
              double q = P.metric[(size_t)(l * P.ndim + m) * P.N + id];
              t += q * q;
            }
            double cs = P.hetmat ? sqrt(P.mat[2 * P.N + id] / P.mat[id]) : B.cs;
            t = cs * sqrt(t) / B.dx[l];
            if (t > dxmax) dxmax = t;
          }
        }
    double test = 1. / dxmax;
    if (test > 1.e-14 && (dxmin <= 1.e-14 || test < dxmin)) dxmin = test;
  }
  if (dxmin <= 1.e-14) throw Err("Error in domain.cpp get_min_dx -- a process does not have any grid spacing values");
  return dxmin;
}

static void set_time_step(Problem& P, bool quiet) {  // problem.cpp:94-134
  double dx = get_min_dx(P);
  if ((P.ttot > 0. && P.nt > 0) && (P.dt == 0 && P.cfl == 0.)) {
    P.dt = P.ttot / (double)P.nt;
    P.cfl = P.dt / dx;
  } else {
    if (P.dt > 0.) {
      if (P.cfl > 0. && !quiet) printf("Cannot specify both dt and cfl, defaulting to dt\n");
      P.cfl = P.dt / dx;
    } else {
      P.dt = P.cfl * dx;
    }
    if (P.ttot > 0.) {
      if (P.nt > 0 && !quiet) printf("Cannot specify both ttot and nt with one of cfl or dt, defaulting to ttot\n");
      P.nt = (int)ceil(P.ttot / P.dt);
      P.ttot = (double)P.nt * P.dt;
    } else {
      P.ttot = (double)P.nt * P.dt;
    }
  }
  if (P.cfl > 1. && !quiet) printf("Warning: CFL ratio > 1, numerical instability is likely\n");
}

// ---------------------------------------------------------------------------
// the hot path

// per-direction material constants, block.cpp:1054-1055,1166-1169,1883-1886
struct DirK {
  double invrho, g2lam, g, lam;
};

static void dir_consts(const Problem& P, const Block& B, double dt, size_t id, DirK K[3]) {
  if (P.hetmat) {
    for (int d = 0; d < P.ndim; d++) {
      double rho = P.mat[id];
      if (P.ndim == 2 && P.mode == 3) {
        double G = P.mat[P.N + id];
        K[d].invrho = dt / rho / B.dx[d];
        K[d].g = dt * G / B.dx[d];
        K[d].g2lam = K[d].lam = 0.;
      } else {
        double lam = P.mat[P.N + id], G = P.mat[2 * P.N + id];
        K[d].invrho = dt / rho / B.dx[d];
        K[d].g2lam = dt * (2. * G + lam) / B.dx[d];
        K[d].g = dt * G / B.dx[d];
        K[d].lam = dt * lam / B.dx[d];
      }
    }
    return;
  }
  K[0].invrho = dt / B.rho / B.dx[0];
  K[0].g2lam = dt * (2. * B.G + B.lam) / B.dx[0];
  K[0].g = dt * B.G / B.dx[0];
  K[0].lam = dt * B.lam / B.dx[0];
  for (int d = 1; d < P.ndim; d++) {
    double r = B.dx[d - 1] / B.dx[d];
    K[d].invrho = K[d - 1].invrho * r;
    K[d].g2lam = K[d - 1].g2lam * r;
    K[d].g = K[d - 1].g * r;
    K[d].lam = K[d - 1].lam * r;
  }
}

// one stencil tap of block::calc_df_* for direction d: adds the
// contribution of neighbour i2 (coefficient c, sign sg) to the accumulators
// of point i1.  Association follows block.cpp:1067-1081 (mode II),
// 1402-1405 (mode III), 1563-1594 (3D).
template <int M>
static inline void tap(double* acc, const Problem& P, size_t i1, size_t i2, double c, double sg, int d,
                       const DirK& K, double invjac) {
  const size_t N = P.N;
  const double* f = P.f.data();
  const double* mt = P.metric.data();
  const double J2 = P.jac[i2];
  if (M == 0) {  // 2D mode II: vx vy sxx sxy syy
    const double m0 = mt[(size_t)(d * 2 + 0) * N + i2], m1 = mt[(size_t)(d * 2 + 1) * N + i2];
    const double n0 = mt[(size_t)(d * 2 + 0) * N + i1], n1 = mt[(size_t)(d * 2 + 1) * N + i1];
    const double vx = f[i2], vy = f[N + i2], sxx = f[2 * N + i2], sxy = f[3 * N + i2], syy = f[4 * N + i2];
    double t0 = (invjac * c * J2 * (m0 * sxx + m1 * sxy));
    double t1 = (invjac * c * J2 * (m0 * sxy + m1 * syy));
    double t2 = c * (K.g2lam * n0 * vx + K.lam * n1 * vy);
    double t3 = K.g * c * (n0 * vy + n1 * vx);
    double t4 = c * (K.g2lam * n1 * vy + K.lam * n0 * vx);
    if (sg > 0.) { acc[0] += t0; acc[1] += t1; acc[2] += t2; acc[3] += t3; acc[4] += t4; }
    else { acc[0] -= t0; acc[1] -= t1; acc[2] -= t2; acc[3] -= t3; acc[4] -= t4; }
  } else if (M == 1) {  // 2D mode III: vz sxz syz
    const double m0 = mt[(size_t)(d * 2 + 0) * N + i2], m1 = mt[(size_t)(d * 2 + 1) * N + i2];
    const double n0 = mt[(size_t)(d * 2 + 0) * N + i1], n1 = mt[(size_t)(d * 2 + 1) * N + i1];
    const double vz = f[i2], sxz = f[N + i2], syz = f[2 * N + i2];
    double t0 = invjac * c * J2 * (m0 * sxz + m1 * syz);
    double t1 = K.g * n0 * c * vz;
    double t2 = K.g * n1 * c * vz;
    if (sg > 0.) { acc[0] += t0; acc[1] += t1; acc[2] += t2; }
    else { acc[0] -= t0; acc[1] -= t1; acc[2] -= t2; }
  } else {  // 3D: vx vy vz sxx sxy sxz syy syz szz
    const double m0 = mt[(size_t)(d * 3 + 0) * N + i2], m1 = mt[(size_t)(d * 3 + 1) * N + i2],
                 m2 = mt[(size_t)(d * 3 + 2) * N + i2];
    const double n0 = mt[(size_t)(d * 3 + 0) * N + i1], n1 = mt[(size_t)(d * 3 + 1) * N + i1],
                 n2 = mt[(size_t)(d * 3 + 2) * N + i1];
    const double vx = f[i2], vy = f[N + i2], vz = f[2 * N + i2];
    const double sxx = f[3 * N + i2], sxy = f[4 * N + i2], sxz = f[5 * N + i2], syy = f[6 * N + i2],
                 syz = f[7 * N + i2], szz = f[8 * N + i2];
    double t0 = (invjac * c * J2 * (m0 * sxx + m1 * sxy + m2 * sxz));
    double t1 = (invjac * c * J2 * (m0 * sxy + m1 * syy + m2 * syz));
    double t2 = (invjac * c * J2 * (m0 * sxz + m1 * syz + m2 * szz));
    double t3 = c * (K.g2lam * n0 * vx + K.lam * (n1 * vy + n2 * vz));
    double t4 = c * K.g * (n0 * vy + n1 * vx);
    double t5 = c * K.g * (n0 * vz + n2 * vx);
    double t6 = c * (K.g2lam * n1 * vy + K.lam * (n0 * vx + n2 * vz));
    double t7 = c * K.g * (n1 * vz + n2 * vy);
    double t8 = c * (K.g2lam * n2 * vz + K.lam * (n0 * vx + n1 * vy));
    if (sg > 0.) {
      acc[0] += t0; acc[1] += t1; acc[2] += t2; acc[3] += t3; acc[4] += t4;
      acc[5] += t5; acc[6] += t6; acc[7] += t7; acc[8] += t8;
    } else {
      acc[0] -= t0; acc[1] -= t1; acc[2] -= t2; acc[3] -= t3; acc[4] -= t4;
      acc[5] -= t5; acc[6] -= t6; acc[7] -= t7; acc[8] -= t8;
    }
  }
}

// block::calc_df for one block (block.cpp:405-427), accumulated point by
// point: for each point the x taps, then y taps, then z taps, each tap adding
// all field terms then the dissipation terms, exactly the per-point order of
// the reference's loop nests.
template <int M>
static void calc_df_block(Problem& P, const Block& B, double dt) {
  const int nfe = (M == 0) ? 5 : (M == 1 ? 3 : 9);
  const int s = P.fd.s;
  const size_t N = P.N;
  const int nd = P.ndim;
  const size_t str[3] = {(size_t)P.Ny * P.Nz, (size_t)P.Nz, 1};
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = B.mlb[0]; i < B.prb[0]; i++)
    for (int j = B.mlb[1]; j < B.prb[1]; j++)
      for (int k = B.mlb[2]; k < B.prb[2]; k++) {
        const size_t i1 = P.idx(i, j, k);
        DirK K[3];
        dir_consts(P, B, dt, i1, K);
        double acc[9];
        for (int q = 0; q < nfe; q++) acc[q] = P.df[(size_t)q * N + i1];
        const int c3[3] = {i, j, k};
        for (int d = 0; d < nd; d++) {
          const double invjac = K[d].invrho / P.jac[i1];
          const int id = c3[d];
          const size_t base = i1 - (size_t)id * str[d];
          int row, ntap, start, step;
          double sg = 1.;
          if (id < B.mc[d]) { row = id - B.mlb[d] + 1; ntap = 3 * (s - 1); start = B.mlb[d]; step = 1; }
          else if (id < B.mrb[d]) { row = 0; ntap = 2 * s - 1; start = id - s + 1; step = 1; }
          else { row = B.prb[d] - id; ntap = 3 * (s - 1); start = B.prb[d] - 1; step = -1; sg = -1.; }
          for (int n = 0; n < ntap; n++) {
            const size_t i2 = base + (size_t)(start + step * n) * str[d];
            tap<M>(acc, P, i1, i2, P.fd.c[row][n], sg, d, K[d], invjac);
            if (P.dissipation) {
              const double dc = P.fd.d[row][n];
              for (int q = 0; q < nfe; q++)
                acc[q] += P.cdiss * (dc * P.jac[i2] * P.f[(size_t)q * N + i2]) / P.jac[i1];
            }
          }
        }
        for (int q = 0; q < nfe; q++) P.df[(size_t)q * N + i1] = acc[q];
      }
  if (M == 0 && P.plastic) {  // block::calc_df_szz, block.cpp:1280-1295
    double nu = 0.5 * B.lam / (B.lam + B.G);
    for (int i = B.mlb[0]; i < B.prb[0]; i++)
      for (int j = B.mlb[1]; j < B.prb[1]; j++) {
        size_t id = P.idx(i, j, 0);
        if (P.hetmat) nu = 0.5 * P.mat[N + id] / (P.mat[N + id] + P.mat[(size_t)P.Ny * P.Nz + id]);
        P.df[5 * N + id] = nu * (P.df[2 * N + id] + P.df[4 * N + id]);
      }
  }
}

// tangent vectors on external faces: max-component rule, boundary.cpp:316-331
static void ext_tangents(const double nn[3], double t1[3], double t2[3]) {
  if (fabs(nn[0]) > fabs(nn[1]) && fabs(nn[0]) > fabs(nn[2])) {
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
}

static BF gather(const Problem& P, size_t id) {  // field -> boundfields, boundary.cpp:333-373
  const size_t N = P.N;
  BF b{};
  if (P.ndim == 3) {
    b.v1 = P.f[id]; b.v2 = P.f[N + id]; b.v3 = P.f[2 * N + id];
    b.s11 = P.f[3 * N + id]; b.s12 = P.f[4 * N + id]; b.s13 = P.f[5 * N + id];
    b.s22 = P.f[6 * N + id]; b.s23 = P.f[7 * N + id]; b.s33 = P.f[8 * N + id];
  } else if (P.mode == 2) {
    b.v1 = P.f[id]; b.v2 = P.f[N + id]; b.v3 = 0.;
    b.s11 = P.f[2 * N + id]; b.s12 = P.f[3 * N + id]; b.s13 = 0.;
    b.s22 = P.f[4 * N + id]; b.s23 = 0.; b.s33 = 0.;
  } else {
    b.v1 = 0.; b.v2 = 0.; b.v3 = P.f[id];
    b.s11 = 0.; b.s12 = 0.; b.s13 = P.f[N + id]; b.s22 = 0.; b.s23 = P.f[2 * N + id]; b.s33 = 0.;
  }
  return b;
}

// boundary::apply_bcs, boundary.cpp:265-492
static void apply_boundary(Problem& P, Bound& bd, double dt) {
  const size_t N = P.N;
  const size_t n = (size_t)bd.n0 * bd.n1;
  for (int i = bd.mlb[0]; i < bd.prb[0]; i++)
    for (int j = bd.mlb[1]; j < bd.prb[1]; j++)
      for (int k = bd.mlb[2]; k < bd.prb[2]; k++) {
        size_t id = P.idx(i, j, k);
        int a, b;
        if (bd.loc < 2) { a = j - bd.mlb[1]; b = k - bd.mlb[2]; }
        else if (bd.loc < 4) { a = i - bd.mlb[0]; b = k - bd.mlb[2]; }
        else { a = i - bd.mlb[0]; b = j - bd.mlb[1]; }
        double nn[3] = {0., 0., 0.}, t1[3], t2[3];
        for (int l = 0; l < P.ndim; l++) nn[l] = bd.nrm[l * n + a * bd.n1 + b];
        double h = dt * bd.dl[a * bd.n1 + b];
        if (P.hetmat) {
          if (P.ndim == 2 && P.mode == 3) {
            bd.cs = sqrt(P.mat[N + id] / P.mat[id]);
            bd.zs = P.mat[id] * bd.cs;
          } else {
            bd.cp = sqrt((P.mat[N + id] + 2. * P.mat[2 * N + id]) / P.mat[id]);
            bd.cs = sqrt(P.mat[2 * N + id] / P.mat[id]);
            bd.zp = P.mat[id] * bd.cp;
            bd.zs = P.mat[id] * bd.cs;
            double rr = bd.cs / bd.cp;
            bd.gamma = 1. - 2. * (rr * rr);
          }
        }
        ext_tangents(nn, t1, t2);
        BF bb = gather(P, id);
        BF br = rotate_xy_nt(bb, nn, t1, t2);
        BF brs = br;
        auto hat = [&](double v, double s, double z, double& vh, double& sh) {  // calc_hat, boundary.cpp:494-503
          sh = (bd.r + 1.) * 0.5 * (s - z * v);
          vh = (bd.r - 1.) * 0.5 * (s / z - v);
        };
        double pv, ps, s1v, s1s, s2v, s2s;
        hat(br.v1, br.s11, bd.zp, pv, ps);
        hat(br.v2, br.s12, bd.zs, s1v, s1s);
        hat(br.v3, br.s13, bd.zs, s2v, s2s);
        br.v1 -= pv;
        br.v2 = 0.;
        br.v3 = 0.;
        br.s22 = bd.gamma * (br.s11 - ps);
        br.s33 = bd.gamma * (br.s11 - ps);
        br.s11 -= ps;
        br.s12 = 0.;
        br.s13 = 0.;
        br.s23 = 0.;
        BF o = rotate_nt_xy(br, nn, t1, t2);
        double* df = P.df.data();
        if (P.ndim == 3) {
          df[id] -= bd.alpha * bd.cp * h * o.v1;
          df[N + id] -= bd.alpha * bd.cp * h * o.v2;
          df[2 * N + id] -= bd.alpha * bd.cp * h * o.v3;
          df[3 * N + id] -= bd.theta * bd.cp * h * o.s11;
          df[4 * N + id] -= bd.theta * bd.cp * h * o.s12;
          df[5 * N + id] -= bd.theta * bd.cp * h * o.s13;
          df[6 * N + id] -= bd.theta * bd.cp * h * o.s22;
          df[7 * N + id] -= bd.theta * bd.cp * h * o.s23;
          df[8 * N + id] -= bd.theta * bd.cp * h * o.s33;
        } else if (P.mode == 2) {
          df[id] -= bd.cp * h * o.v1;
          df[N + id] -= bd.cp * h * o.v2;
          df[2 * N + id] -= bd.cp * h * o.s11;
          df[3 * N + id] -= bd.cp * h * o.s12;
          df[4 * N + id] -= bd.cp * h * o.s22;
          if (P.plastic) df[5 * N + id] -= bd.cp * h * o.s33;
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
        o = rotate_nt_xy(brs, nn, t1, t2);
        if (P.ndim == 3) {
          df[id] -= bd.alpha * bd.cs * h * o.v1;
          df[N + id] -= bd.alpha * bd.cs * h * o.v2;
          df[2 * N + id] -= bd.alpha * bd.cs * h * o.v3;
          df[3 * N + id] -= bd.theta * bd.cs * h * o.s11;
          df[4 * N + id] -= bd.theta * bd.cs * h * o.s12;
          df[5 * N + id] -= bd.theta * bd.cs * h * o.s13;
          df[6 * N + id] -= bd.theta * bd.cs * h * o.s22;
          df[7 * N + id] -= bd.theta * bd.cs * h * o.s23;
          df[8 * N + id] -= bd.theta * bd.cs * h * o.s33;
        } else if (P.mode == 2) {
          df[id] -= bd.cs * h * o.v1;
          df[N + id] -= bd.cs * h * o.v2;
          df[2 * N + id] -= bd.cs * h * o.s11;
          df[3 * N + id] -= bd.cs * h * o.s12;
          df[4 * N + id] -= bd.cs * h * o.s22;
        } else {
          df[id] -= bd.cs * h * o.v3;
          df[N + id] -= bd.cs * h * o.s13;
          df[2 * N + id] -= bd.cs * h * o.s23;
        }
      }
}

struct IfChar {
  double v1, v2, s1, s2;
};
struct IfHat {
  double v11, v12, v13, s11, s12, s13, v21, v22, v23, s21, s22, s23;
};

static IfChar solve_locked(const IfChar& c, double z1, double z2) {  // interface.cpp:1017-1028
  IfChar o;
  o.s1 = (z2 * (c.s1 - z1 * c.v1) + z1 * (c.s2 + z2 * c.v2)) / (z1 + z2);
  o.s2 = o.s1;
  o.v1 = (o.s1 - c.s1) / z1 + c.v1;
  o.v2 = o.v1;
  return o;
}

// slipweak::calc_mu, slipweak.cpp:253-308
static double sw_mu_value(double u, double t, double dct, double must, double mudt, double trupt, double tct) {
  double f1, f2;
  if (dct == 0. || u >= dct) f1 = 1.;
  else f1 = u / dct;
  if (trupt <= 0. || t < trupt) f2 = 0.;
  else if (t >= trupt && t < trupt + tct) f2 = (tct == 0.) ? 1. : (t - trupt) / tct;
  else f2 = 1.;
  return (f1 >= f2) ? must + (mudt - must) * f1 : must + (mudt - must) * f2;
}

// slipweak::solve_fs for one point, slipweak.cpp:206-251: compressive normal
// stress -> strength c0 + mu |sn|, locked if it exceeds phi; tensile -> slip
// once phi exceeds the cohesion alone
static void sw_solve_point(double phi, double eta, double snc, double c0t, double mu, double& V, double& S) {
  if (snc < 0.) {
    if (c0t + mu * fabs(snc) > phi) { V = 0.; S = phi; }
    else { S = c0t + mu * fabs(snc); V = (phi - S) / eta; }
  } else {
    if (phi > c0t) { S = c0t; V = (phi - S) / eta; }
    else { V = 0.; S = phi; }
  }
}

// stz::calc_dstatedt + chihat, stz.cpp:311-356
static double stz_dstatedt(double V, double S, double chi, double c0t, double Rt, double betat, double chiwt,
                           double v1t) {
  if (V > v1t) throw Err("Slip velocity exceeds melting value in stz.cpp");
  double chihat = (V == 0.) ? chiwt : chiwt / log(v1t / V);
  return V * S / c0t * (1. - chi / chihat) - Rt * exp(-betat / chi);
}

// friction law for one point: (phi, eta, snc) -> (V, S); slipweak.cpp:206-251,
// friction.cpp:400-435, stz.cpp:263-356
static void solve_fs(Iface& I, int i, int j, double phi, double eta, double snc, double t, double& V, double& S) {
  const int index = i * I.n[1] + j;
  if (I.type == 2) {
    double c0t = 0.;
    for (const Pert& p : I.perts) c0t += p.get(3, i, j, t);
    if (I.param_file) c0t += I.prm[3][index];
    if (snc < 0.) {
      double dct = 0., must = 0., mudt = 0., trupt = 0., tct = 0.;
      for (const Pert& p : I.perts) {
        dct += p.get(0, i, j, t);
        must += p.get(1, i, j, t);
        mudt += p.get(2, i, j, t);
        trupt += p.get(4, i, j, t);
        tct += p.get(5, i, j, t);
      }
      if (I.param_file) {
        dct += I.prm[0][index];
        must += I.prm[1][index];
        mudt += I.prm[2][index];
        trupt += I.prm[4][index];
        tct += I.prm[5][index];
      }
      double mu = sw_mu_value(I.u[index], t, dct, must, mudt, trupt, tct);
      sw_solve_point(phi, eta, snc, c0t, mu, V, S);
    } else {
      sw_solve_point(phi, eta, snc, c0t, 0., V, S);
    }
    return;
  }
  // base friction::solve_fs with calc_mu of frictionless (0) or stz
  if (snc < 0.) {
    double mu = 0.;
    if (I.type == 3) {
      double prm[8] = {phi, eta, snc, I.state[index], 0., 0., 0., 0.};
      for (const Pert& p : I.perts) {
        prm[4] += p.get(0, i, j, t);
        prm[5] += p.get(1, i, j, t);
        prm[6] += p.get(2, i, j, t);
        prm[7] += p.get(3, i, j, t);
      }
      if (I.param_file) {
        prm[4] += I.prm[0][index];
        prm[5] += I.prm[1][index];
        prm[6] += I.prm[2][index];
        prm[7] += I.prm[3][index];
      }
      if (prm[7] >= -phi / snc) mu = -phi / snc;
      else mu = solve_newton(prm[7], -phi / snc, prm);
    }
    if (mu * fabs(snc) > phi) { V = 0.; S = phi; }
    else { S = mu * fabs(snc); V = (phi - S) / eta; }
  } else {
    S = 0.;
    V = phi / eta;
  }
  if (I.has_state) {  // stz::calc_dstatedt, stz.cpp:311-356
    double c0t = 0., Rt = 0., betat = 0., chiwt = 0., v1t = 0.;
    for (const Pert& p : I.perts) {
      c0t += p.get(4, i, j, t);
      Rt += p.get(5, i, j, t);
      betat += p.get(6, i, j, t);
      chiwt += p.get(7, i, j, t);
      v1t += p.get(8, i, j, t);
    }
    if (I.param_file) {
      c0t += I.prm[4][index];
      Rt += I.prm[5][index];
      betat += I.prm[6][index];
      chiwt += I.prm[7][index];
      v1t += I.prm[8][index];
    }
    I.dstatedt[index] = stz_dstatedt(V, S, I.state[index], c0t, Rt, betat, chiwt, v1t);
  }
}

// interface::solve_interface + friction::solve_interface/solve_friction,
// interface.cpp:972-1015, friction.cpp:259-398
static IfHat solve_interface(Problem& P, Iface& I, const BF& b1, const BF& b2, double zp1, double zs1, double zp2,
                             double zs2, int i, int j, double t) {
  IfHat o;
  IfChar cp{b1.v1, b2.v1, b1.s11, b2.s11};
  IfChar hp = solve_locked(cp, zp1, zp2);
  if (I.type == 0) {
    IfChar c1{b1.v2, b2.v2, b1.s12, b2.s12};
    IfChar h1 = solve_locked(c1, zs1, zs2);
    IfChar c2{b1.v3, b2.v3, b1.s13, b2.s13};
    IfChar h2 = solve_locked(c2, zs1, zs2);
    o.v11 = hp.v1; o.v21 = hp.v2; o.s11 = hp.s1; o.s21 = hp.s2;
    o.v12 = h1.v1; o.v22 = h1.v2; o.s12 = h1.s1; o.s22 = h1.s2;
    o.v13 = h2.v1; o.v23 = h2.v2; o.s13 = h2.s1; o.s23 = h2.s2;
    return o;
  }
  double snc = hp.s1;
  double v12 = b1.v2, v22 = b2.v2, v13 = b1.v3, v23 = b2.v3;
  double s12 = b1.s12, s22 = b2.s12, s13 = b1.s13, s23 = b2.s13;
  const double z1 = zs1, z2 = zs2;
  const double eta = z1 * z2 / (z1 + z2);
  const int index = i * I.n[1] + j;
  if (I.load_file) {
    snc += I.ls1[index];
    s12 += I.ls2[index];
    s22 += I.ls2[index];
    s13 += I.ls3[index];
    s23 += I.ls3[index];
  }
  for (const Pert& L : I.loads) {
    snc += L.get(0, i, j, t);
    s12 += L.get(1, i, j, t);
    s22 += L.get(1, i, j, t);
    s13 += L.get(2, i, j, t);
    s23 += L.get(2, i, j, t);
  }
  double phi2 = eta * (s12 / z1 - v12 + s22 / z2 + v22);
  double phi3 = eta * (s13 / z1 - v13 + s23 / z2 + v23);
  double phi = sqrt(phi2 * phi2 + phi3 * phi3);
  double V, S;
  solve_fs(I, i, j, phi, eta, snc, t, V, S);
  double vv2, vv3;
  if (V == 0.) { vv2 = 0.; vv3 = 0.; }
  else { vv2 = V * phi2 / (eta * V + S); vv3 = V * phi3 / (eta * V + S); }
  o.s12 = phi2 - eta * vv2;
  o.s22 = o.s12;
  o.s13 = phi3 - eta * vv3;
  o.s23 = o.s13;
  o.v12 = (o.s12 - s12) / z1 + v12;
  o.v22 = (-o.s22 + s22) / z2 + v22;
  o.v13 = (o.s13 - s13) / z1 + v13;
  o.v23 = (-o.s23 + s23) / z2 + v23;
  const size_t nn = (size_t)I.n[0] * I.n[1];
  I.v[index] = V;
  I.s[index] = S;
  I.sn[index] = snc;
  if (P.ndim == 3) {
    I.vx[index] = vv2; I.vx[nn + index] = vv3;
    I.sx[index] = o.s12; I.sx[nn + index] = o.s13;
  } else if (P.mode == 2) {
    I.vx[index] = vv2; I.sx[index] = o.s12;
  } else {
    I.vx[index] = vv3; I.sx[index] = o.s13;
  }
  if (I.load_file) {
    o.s12 -= I.ls2[index];
    o.s22 -= I.ls2[index];
    o.s13 -= I.ls3[index];
    o.s23 -= I.ls3[index];
  }
  for (const Pert& L : I.loads) {
    o.s12 -= L.get(1, i, j, t);
    o.s22 -= L.get(1, i, j, t);
    o.s13 -= L.get(2, i, j, t);
    o.s23 -= L.get(2, i, j, t);
  }
  o.v11 = hp.v1; o.v21 = hp.v2; o.s11 = hp.s1; o.s21 = hp.s2;
  return o;
}

// interface::apply_bcs, interface.cpp:573-970
static void apply_interface(Problem& P, Iface& I, double dt, double t, bool no_sat) {
  if (!I.is_friction && no_sat) return;
  const size_t N = P.N;
  const size_t n = (size_t)I.n[0] * I.n[1];
  const double* s0 = P.s0;
  for (int i = I.mlb[0]; i < I.prb[0]; i++)
    for (int j = I.mlb[1]; j < I.prb[1]; j++)
      for (int k = I.mlb[2]; k < I.prb[2]; k++) {
        size_t id1 = P.idx(i, j, k);
        size_t id2 = P.idx(i + I.delta[0], j + I.delta[1], k + I.delta[2]);
        int ii, jj;
        if (I.dir == 0) { ii = j - I.mlb[1]; jj = k - I.mlb[2]; }
        else if (I.dir == 1) { ii = i - I.mlb[0]; jj = k - I.mlb[2]; }
        else { ii = i - I.mlb[0]; jj = j - I.mlb[1]; }
        double h1 = dt * I.dl1[ii * I.n[1] + jj];
        double h2 = dt * I.dl2[ii * I.n[1] + jj];
        if (P.hetmat) {
          if (P.ndim == 2 && P.mode == 3) {
            I.cs1 = sqrt(P.mat[N + id1] / P.mat[id1]); I.zs1 = P.mat[id1] * I.cs1;
            I.cs2 = sqrt(P.mat[N + id2] / P.mat[id2]); I.zs2 = P.mat[id2] * I.cs2;
          } else {
            I.cp1 = sqrt((P.mat[N + id1] + 2. * P.mat[2 * N + id1]) / P.mat[id1]);
            I.cs1 = sqrt(P.mat[2 * N + id1] / P.mat[id1]);
            I.zp1 = P.mat[id1] * I.cp1; I.zs1 = P.mat[id1] * I.cs1;
            double r1 = I.cs1 / I.cp1; I.gamma1 = 1. - 2. * (r1 * r1);
            I.cp2 = sqrt((P.mat[N + id2] + 2. * P.mat[2 * N + id2]) / P.mat[id2]);
            I.cs2 = sqrt(P.mat[2 * N + id2] / P.mat[id2]);
            I.zp2 = P.mat[id2] * I.cp2; I.zs2 = P.mat[id2] * I.cs2;
            double r2 = I.cs2 / I.cp2; I.gamma2 = 1. - 2. * (r2 * r2);
          }
        }
        double nn[3] = {0., 0., 0.}, t1[3], t2[3];
        for (int l = 0; l < P.ndim; l++) nn[l] = I.nrm[l * n + ii * I.n[1] + jj];
        if (I.dir == 0) {
          t1[2] = 0.;
          t1[1] = nn[0] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
          t1[0] = -nn[1] / sqrt(nn[0] * nn[0] + nn[1] * nn[1]);
        } else if (I.dir == 1) {
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
        if (I.dir == 1) { t2[0] = -t2[0]; t2[1] = -t2[1]; t2[2] = -t2[2]; }
        BF b1, b2;
        if (P.ndim == 3) {
          b1 = gather(P, id1);
          b2 = gather(P, id2);
          b1.s11 += s0[0]; b1.s12 += s0[1]; b1.s13 += s0[2]; b1.s22 += s0[3]; b1.s23 += s0[4]; b1.s33 += s0[5];
          b2.s11 += s0[0]; b2.s12 += s0[1]; b2.s13 += s0[2]; b2.s22 += s0[3]; b2.s23 += s0[4]; b2.s33 += s0[5];
          if (P.hetstress) {
            const double* sh = P.s.data();
            b1.s11 += sh[id1]; b1.s12 += sh[N + id1]; b1.s13 += sh[2 * N + id1];
            b1.s22 += sh[3 * N + id1]; b1.s23 += sh[4 * N + id1]; b1.s33 += sh[5 * N + id1];
            b2.s11 += sh[id2]; b2.s12 += sh[N + id2]; b2.s13 += sh[2 * N + id2];
            b2.s22 += sh[3 * N + id2]; b2.s23 += sh[4 * N + id2]; b2.s33 += sh[5 * N + id2];
          }
        } else if (P.mode == 2) {
          b1 = gather(P, id1);
          b2 = gather(P, id2);
          b1.s11 += s0[0]; b1.s12 += s0[1]; b1.s22 += s0[3];
          b2.s11 += s0[0]; b2.s12 += s0[1]; b2.s22 += s0[3];
          if (P.hetstress) {
            const double* sh = P.s.data();
            b1.s11 += sh[id1]; b1.s12 += sh[N + id1]; b1.s22 += sh[2 * N + id1];
            b2.s11 += sh[id2]; b2.s12 += sh[N + id2]; b2.s22 += sh[2 * N + id2];
          }
        } else {
          b1 = gather(P, id1);
          b2 = gather(P, id2);
          b1.s11 = s0[0]; b1.s13 += s0[2]; b1.s22 = s0[3]; b1.s23 += s0[4];
          b2.s11 = s0[0]; b2.s13 += s0[2]; b2.s22 = s0[3]; b2.s23 += s0[4];
          if (P.hetstress) {
            const double* sh = P.s.data();
            b1.s13 += sh[id1]; b1.s23 += sh[N + id1];
            b2.s13 += sh[id2]; b2.s23 += sh[N + id2];
          }
        }
        BF r1 = rotate_xy_nt(b1, nn, t1, t2);
        BF r2 = rotate_xy_nt(b2, nn, t1, t2);
        BF rs1 = r1, rs2 = r2;
        IfHat hat = solve_interface(P, I, r1, r2, I.zp1, I.zs1, I.zp2, I.zs2, ii, jj, t);
        if (no_sat) continue;
        r1.v1 -= hat.v11; r1.v2 = 0.; r1.v3 = 0.;
        r1.s22 = I.gamma1 * (r1.s11 - hat.s11);
        r1.s33 = I.gamma1 * (r1.s11 - hat.s11);
        r1.s11 -= hat.s11; r1.s12 = 0.; r1.s13 = 0.; r1.s23 = 0.;
        r2.v1 -= hat.v21; r2.v2 = 0.; r2.v3 = 0.;
        r2.s22 = I.gamma2 * (r2.s11 - hat.s21);
        r2.s33 = I.gamma2 * (r2.s11 - hat.s21);
        r2.s11 -= hat.s21; r2.s12 = 0.; r2.s13 = 0.; r2.s23 = 0.;
        BF o1 = rotate_nt_xy(r1, nn, t1, t2);
        BF o2 = rotate_nt_xy(r2, nn, t1, t2);
        double* df = P.df.data();
        if (P.ndim == 3) {
          const double* a1 = &o1.v1;
          const double* a2 = &o2.v1;
          for (int q = 0; q < 9; q++) df[(size_t)q * N + id1] -= I.cp1 * h1 * a1[q];
          for (int q = 0; q < 9; q++) df[(size_t)q * N + id2] -= I.cp2 * h2 * a2[q];
        } else if (P.mode == 2) {
          df[id1] -= I.cp1 * h1 * o1.v1;
          df[N + id1] -= I.cp1 * h1 * o1.v2;
          df[2 * N + id1] -= I.cp1 * h1 * o1.s11;
          df[3 * N + id1] -= I.cp1 * h1 * o1.s12;
          df[4 * N + id1] -= I.cp1 * h1 * o1.s22;
          if (P.plastic) df[5 * N + id1] -= I.cp1 * h1 * o1.s33;
          df[id2] -= I.cp2 * h2 * o2.v1;
          df[N + id2] -= I.cp2 * h2 * o2.v2;
          df[2 * N + id2] -= I.cp2 * h2 * o2.s11;
          df[3 * N + id2] -= I.cp2 * h2 * o2.s12;
          df[4 * N + id2] -= I.cp2 * h2 * o2.s22;
          if (P.plastic) df[5 * N + id2] -= I.cp2 * h2 * o2.s33;
        }
        rs1.v1 = 0.; rs1.v2 -= hat.v12; rs1.v3 -= hat.v13;
        rs1.s11 = 0.; rs1.s12 -= hat.s12; rs1.s13 -= hat.s13; rs1.s22 = 0.; rs1.s23 = 0.; rs1.s33 = 0.;
        rs2.v1 = 0.; rs2.v2 -= hat.v22; rs2.v3 -= hat.v23;
        rs2.s11 = 0.; rs2.s12 -= hat.s22; rs2.s13 -= hat.s23; rs2.s22 = 0.; rs2.s23 = 0.; rs2.s33 = 0.;
        o1 = rotate_nt_xy(rs1, nn, t1, t2);
        o2 = rotate_nt_xy(rs2, nn, t1, t2);
        if (P.ndim == 3) {
          const double* a1 = &o1.v1;
          const double* a2 = &o2.v1;
          for (int q = 0; q < 9; q++) df[(size_t)q * N + id1] -= I.cs1 * h1 * a1[q];
          for (int q = 0; q < 9; q++) df[(size_t)q * N + id2] -= I.cs2 * h2 * a2[q];
        } else if (P.mode == 2) {
          df[id1] -= I.cs1 * h1 * o1.v1;
          df[N + id1] -= I.cs1 * h1 * o1.v2;
          df[2 * N + id1] -= I.cs1 * h1 * o1.s11;
          df[3 * N + id1] -= I.cs1 * h1 * o1.s12;
          df[4 * N + id1] -= I.cs1 * h1 * o1.s22;
          df[id2] -= I.cs2 * h2 * o2.v1;
          df[N + id2] -= I.cs2 * h2 * o2.v2;
          df[2 * N + id2] -= I.cs2 * h2 * o2.s11;
          df[3 * N + id2] -= I.cs2 * h2 * o2.s12;
          df[4 * N + id2] -= I.cs2 * h2 * o2.s22;
        } else {
          df[id1] -= I.cs1 * h1 * o1.v3;
          df[N + id1] -= I.cs1 * h1 * o1.s13;
          df[2 * N + id1] -= I.cs1 * h1 * o1.s23;
          df[id2] -= I.cs2 * h2 * o2.v3;
          df[N + id2] -= I.cs2 * h2 * o2.s13;
          df[2 * N + id2] -= I.cs2 * h2 * o2.s23;
        }
      }
}

// fields::set_stress / remove_stress, fields.cpp:184-220
static void set_stress(Problem& P, double sign) {
  const size_t N = P.N;
  for (int i = 0; i < P.ns; i++) {
    double s0 = P.s0[P.sidx[i]];
    double* f = P.f.data() + (size_t)(P.nv + i) * N;
    if (sign > 0.) for (size_t j = 0; j < N; j++) f[j] += s0;
    else for (size_t j = 0; j < N; j++) f[j] -= s0;
  }
  if (!P.hetstress) return;
  for (int i = 0; i < P.ns; i++) {
    double* f = P.f.data() + (size_t)(P.nv + i) * N;
    const double* s = P.s.data() + (size_t)i * N;
    if (sign > 0.) for (size_t j = 0; j < N; j++) f[j] += s[j];
    else for (size_t j = 0; j < N; j++) f[j] -= s[j];
  }
}

// block::calc_plastic, block.cpp:442-589
static void calc_plastic(Problem& P, const Block& B, double dt) {
  const size_t N = P.N;
  double* f = P.f.data();
  const bool pt = P.plastic_tensor;
  for (int i = B.mlb[0]; i < B.prb[0]; i++)
    for (int j = B.mlb[1]; j < B.prb[1]; j++)
      for (int k = B.mlb[2]; k < B.prb[2]; k++) {
        size_t id = P.idx(i, j, k);
        PS si{};
        int base;  // first plastic field index
        if (P.ndim == 3) {
          si.sxx = f[3 * N + id]; si.sxy = f[4 * N + id]; si.sxz = f[5 * N + id];
          si.syy = f[6 * N + id]; si.syz = f[7 * N + id]; si.szz = f[8 * N + id];
          base = 9;
        } else if (P.mode == 2) {
          si.sxx = f[2 * N + id]; si.sxy = f[3 * N + id]; si.sxz = 0.;
          si.syy = f[4 * N + id]; si.syz = 0.; si.szz = f[5 * N + id];
          base = 6;
        } else {
          si.sxx = P.s0[0]; si.sxy = 0.; si.sxz = f[1 * N + id];
          si.syy = P.s0[3]; si.syz = f[2 * N + id]; si.szz = P.s0[5];
          base = 3;
        }
        si.lambda = f[(size_t)base * N + id];
        si.gammap = f[(size_t)(base + 1) * N + id];
        if (pt) {
          si.epxx = f[(size_t)(base + 2) * N + id]; si.epxy = f[(size_t)(base + 3) * N + id];
          si.epxz = f[(size_t)(base + 4) * N + id]; si.epyy = f[(size_t)(base + 5) * N + id];
          si.epyz = f[(size_t)(base + 6) * N + id]; si.epzz = f[(size_t)(base + 7) * N + id];
        }
        double kk, g;
        if (P.hetmat) {
          kk = P.mat[N + id] + 2. / 3. * P.mat[2 * N + id];
          g = P.mat[2 * N + id];
        } else {
          kk = B.lam + 2. / 3. * B.G;
          g = B.G;
        }
        PS so = plastic_flow(dt, si, kk, g, B.pm);
        if (P.ndim == 3) {
          f[3 * N + id] = so.sxx; f[4 * N + id] = so.sxy; f[5 * N + id] = so.sxz;
          f[6 * N + id] = so.syy; f[7 * N + id] = so.syz; f[8 * N + id] = so.szz;
        } else if (P.mode == 2) {
          f[2 * N + id] = so.sxx; f[3 * N + id] = so.sxy; f[4 * N + id] = so.syy; f[5 * N + id] = so.szz;
        } else {
          f[1 * N + id] = so.sxz; f[2 * N + id] = so.syz;
        }
        f[(size_t)base * N + id] = so.lambda;
        f[(size_t)(base + 1) * N + id] = so.gammap;
        if (pt) {
          f[(size_t)(base + 2) * N + id] = so.epxx; f[(size_t)(base + 3) * N + id] = so.epxy;
          f[(size_t)(base + 4) * N + id] = so.epxz; f[(size_t)(base + 5) * N + id] = so.epyy;
          f[(size_t)(base + 6) * N + id] = so.epyz; f[(size_t)(base + 7) * N + id] = so.epzz;
        }
      }
}

// domain::do_rk_stage, domain.cpp:228-305
static void do_rk_stage(Problem& P, int stage, double t) {
  const double dt = P.dt;
  const double A = P.rk.A[stage], Bc = P.rk.B[stage];
  for (double& v : P.df) v *= A;
  for (Iface& I : P.ifaces) {  // friction::scale_df, friction.cpp:444-463 (dstate over its own size)
    if (!I.is_friction) continue;
    for (double& v : I.dux) v *= A;
    for (double& v : I.du) v *= A;
    if (I.has_state) for (double& v : I.dstate) v *= A;
  }
  for (Block& B : P.blocks) {
    if (P.ndim == 3) calc_df_block<2>(P, B, dt);
    else if (P.mode == 2) calc_df_block<0>(P, B, dt);
    else calc_df_block<1>(P, B, dt);
    for (int loc = 0; loc < 2 * P.ndim; loc++)
      if (B.bnd[loc].active) apply_boundary(P, B.bnd[loc], dt);
  }
  const double ts = t + P.rk.C[stage] * dt;
  for (Iface& I : P.ifaces) apply_interface(P, I, dt, ts, false);
  for (Iface& I : P.ifaces) {  // friction::calc_df, friction.cpp:465-484
    if (!I.is_friction) continue;
    for (size_t q = 0; q < I.dux.size(); q++) I.dux[q] += dt * I.vx[q];
    for (size_t q = 0; q < I.du.size(); q++) I.du[q] += dt * I.v[q];
    if (I.has_state) for (size_t q = 0; q < I.dstate.size(); q++) I.dstate[q] += dt * I.dstatedt[q];
  }
  for (Iface& I : P.ifaces) {  // friction::update, friction.cpp:486-505
    if (!I.is_friction) continue;
    for (size_t q = 0; q < I.ux.size(); q++) I.ux[q] += Bc * I.dux[q];
    for (size_t q = 0; q < I.u.size(); q++) I.u[q] += Bc * I.du[q];
    if (I.has_state) for (size_t q = 0; q < I.state.size(); q++) I.state[q] += Bc * I.dstate[q];
  }
  const size_t nupd = (size_t)P.nf * P.N;
  for (size_t q = 0; q < nupd; q++) P.f[q] += Bc * P.df[q];
  if (stage + 1 == P.rk.nstages && P.plastic) {
    set_stress(P, 1.);
    for (Block& B : P.blocks) calc_plastic(P, B, dt);
    set_stress(P, -1.);
    for (Iface& I : P.ifaces) apply_interface(P, I, dt, ts, true);
  }
}

// ---------------------------------------------------------------------------
// outputs (outputlist.cpp, outputunit.cpp) and fronts (frontlist.cpp, front.cpp)

static char endian_char() {
  union {
    uint32_t i;
    char c[4];
  } u = {0x01020304};
  if (u.c[0] == 1) return '>';
  if (u.c[0] == 4) return '<';
  return '=';
}

static int field_index(const Problem& P, const std::string& f, bool& is_iface) {
  static const char* f3[] = {"vx", "vy", "vz", "sxx", "sxy", "sxz", "syy", "syz", "szz",
                              "lambda", "gammap", "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
  static const char* i3[] = {"Vx", "Vy", "Vz", "V", "Ux", "Uy", "Uz", "U", "Sx", "Sy", "Sz", "S", "Sn", "state"};
  static const char* f2[] = {"vx", "vy", "sxx", "sxy", "syy", "szz", "lambda", "gammap",
                              "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
  static const char* i2[] = {"Vx", "Vy", "V", "Ux", "Uy", "U", "Sx", "Sy", "S", "Sn", "state"};
  static const char* f2m3[] = {"vz", "sxz", "syz", "lambda", "gammap", "epxx", "epxy", "epxz", "epyy", "epyz", "epzz"};
  static const char* i2m3[] = {"Vz", "V", "Uz", "U", "Sz", "S", "Sn", "state"};
  const char** fl;
  const char** il;
  int nfl, nil;
  if (P.ndim == 3) { fl = f3; nfl = 17; il = i3; nil = 14; }
  else if (P.mode == 2) { fl = f2; nfl = 14; il = i2; nil = 11; }
  else { fl = f2m3; nfl = 11; il = i2m3; nil = 8; }
  for (int i = 0; i < nfl; i++) if (f == fl[i]) { is_iface = false; return i; }
  for (int i = 0; i < nil; i++) if (f == il[i]) { is_iface = true; return i; }
  throw Err("Error in specifying output field in outputunit.cpp");
}

static void open_outputs(Problem& P, InText& in) {  // outputlist.cpp:11-75, outputunit.cpp:13-741
  in.rewind();
  if (!in.scan("[fdfault.outputlist]")) throw Err("Error reading outputlist from input file");
  std::string line;
  const char endian = endian_char();
  while (in.getline(line)) {
    if (line.empty()) break;
    OutUnit U;
    U.name = line;
    U.field_name = in.tok();
    int tm = in.integer(), tp = in.integer(), ts = in.integer();
    int xm[3], xp[3], xs[3];
    for (int i = 0; i < 3; i++) { xm[i] = in.integer(); xp[i] = in.integer(); xs[i] = in.integer(); }
    in.getline(line);
    if (ts <= 0 || tp < tm) throw Err("invalid output time range");
    for (int i = 0; i < 3; i++)
      if (xm[i] >= P.nx[i] || xs[i] <= 0 || xp[i] < xm[i]) throw Err("invalid output index range");
    bool is_if = false;
    U.field = field_index(P, U.field_name, is_if);
    const std::string& fn = U.field_name;
    if (!is_if) {
      if ((fn == "lambda" || fn == "gammap") && !P.plastic) throw Err("plastic output requested for elastic problem");
      if (fn.size() == 4 && fn.substr(0, 2) == "ep" && !P.plastic_tensor) throw Err("plastic strain output requires plastic tensor");
      if (P.ndim == 2 && P.mode == 2 && fn == "szz" && !P.plastic) throw Err("szz output requires plasticity in mode 2");
    }
    U.location = -1;
    if (is_if) {
      for (int i = 0; i < P.nifaces; i++) {
        const Iface& I = P.ifaces[i];
        int d = I.dir;
        if (xm[d] != xp[d]) continue;
        bool ok = (xm[d] == I.xm[d] || xm[d] == I.xm[d] - 1);
        for (int q = 0; q < 3; q++)
          if (q != d) ok = ok && xm[q] >= I.xm[q] && xm[q] <= I.xp[q];
        if (ok) { U.location = i; break; }
      }
      if (U.location == -1) throw Err("Could not find interface corresponding to input indices in outputitem.cpp");
      const Iface& I = P.ifaces[U.location];
      if (!I.is_friction) throw Err("Interface for output is not frictional in outputunit.cpp");
      if (fn == "state" && !I.has_state) throw Err("Interface for output does not have a state variable in outputunit.cpp");
    }
    U.tm = tm;
    U.tp = (tp > P.nt) ? P.nt : tp;
    U.tp = U.tp - (U.tp - U.tm) % ts;
    U.ts = ts;
    int ntout = (U.tp - U.tm) / ts + 1;
    for (int i = 0; i < 3; i++) {
      if (U.location == -1) {
        U.xm[i] = (xm[i] < 0) ? 0 : xm[i];
        U.xp[i] = (xp[i] >= P.nx[i]) ? P.nx[i] - 1 : xp[i] - (xp[i] - xm[i]) % xs[i];
      } else {
        const Iface& I = P.ifaces[U.location];
        U.xm[i] = (xm[i] < I.xm[i]) ? I.xm[i] : xm[i];
        U.xp[i] = (xp[i] > I.xp[i]) ? I.xp[i] : xp[i] - (xp[i] - xm[i]) % xs[i];
        if (U.xm[i] > U.xp[i]) U.xp[i] = U.xm[i];
      }
      U.xs[i] = xs[i];
      U.nx[i] = (U.xp[i] - U.xm[i]) / U.xs[i] + 1;
    }
    if (P.ndim == 2) { U.xm[2] = 0; U.xp[2] = 0; U.xs[2] = 1; U.nx[2] = 1; }
    for (int i = 0; i < 3; i++) if (U.nx[i] == 1) U.xs[i] = 1;
    if (U.location >= 0) {  // which interface array and component (outputunit.cpp:403-470)
      const Iface& I = P.ifaces[U.location];
      int d = I.dir, fld = U.field;
      // component index in the vector arrays: tangential axes in order, skipping the normal
      int axis = -1;
      std::string base;
      if (fn.size() == 2 && (fn[0] == 'V' || fn[0] == 'U' || fn[0] == 'S') && fn != "Sn") {
        axis = (P.ndim == 2 && P.mode == 3) ? 2 : fn[1] - 'x';
      }
      (void)fld;
      U.comp = 0;
      if (axis >= 0) {
        if (axis == d) throw Err("invalid interface output component");
        if (P.ndim == 3) {
          int a = (d == 0) ? 1 : 0;
          U.comp = (axis == a) ? 0 : 1;
        } else {
          U.comp = 0;
        }
      }
    }
    std::string path = P.datadir + P.name + "_" + U.name + "_" + fn + ".dat";
    U.fp = fopen(path.c_str(), "wb");
    if (!U.fp) throw Err("Error opening file in outputunit.cpp");
    const char* xyz[3] = {"x", "y", "z"};
    for (int i = 0; i < P.ndim; i++) {
      std::string xpath = P.datadir + P.name + "_" + U.name + "_" + xyz[i] + ".dat";
      FILE* xf = fopen(xpath.c_str(), "wb");
      if (!xf) throw Err("Error opening file in outputunit.cpp");
      for (int a = U.xm[0]; a <= U.xp[0]; a += U.xs[0])
        for (int b = U.xm[1]; b <= U.xp[1]; b += U.xs[1])
          for (int c = U.xm[2]; c <= U.xp[2]; c += U.xs[2]) {
            double v = P.x[(size_t)i * P.N + P.idx(a, b, c)];
            fwrite(&v, sizeof(double), 1, xf);
          }
      fclose(xf);
    }
    if (U.tp > U.tm) {
      U.tfp = fopen((P.datadir + P.name + "_" + U.name + "_t.dat").c_str(), "wb");
      if (!U.tfp) throw Err("Error opening file in outputunit.cpp");
    }
    {
      FILE* m = fopen((P.datadir + P.name + "_" + U.name + ".m").c_str(), "w");
      if (!m) throw Err("Error writing parameters to matlab file");
      fprintf(m, "endian = '%c';\n", endian == '<' ? 'l' : (endian == '>' ? 'b' : 'n'));
      fprintf(m, "field = '%s';\n", fn.c_str());
      fprintf(m, "nt = %d;\nnx = %d;\nny = %d;\nnz = %d;\n", ntout, U.nx[0], U.nx[1], U.nx[2]);
      fclose(m);
      FILE* py = fopen((P.datadir + P.name + "_" + U.name + ".py").c_str(), "w");
      if (!py) throw Err("Error writing parameters to python file");
      fprintf(py, "endian = '%c'\n", endian);
      fprintf(py, "field = '%s';\n", fn.c_str());
      fprintf(py, "nt = %d\nnx = %d\nny = %d\nnz = %d\n", ntout, U.nx[0], U.nx[1], U.nx[2]);
      fclose(py);
    }
    P.outs.push_back(U);
  }
}

static const std::vector<double>* iface_array(const Problem& P, const Iface& I, const std::string& fn) {
  (void)P;
  if (fn == "V") return &I.v;
  if (fn == "U") return &I.u;
  if (fn == "S") return &I.s;
  if (fn == "Sn") return &I.sn;
  if (fn == "state") return &I.state;
  if (fn[0] == 'V') return &I.vx;
  if (fn[0] == 'U') return &I.ux;
  return &I.sx;
}

static void write_outputs(Problem& P, int tstep) {  // outputunit.cpp:771-889
  for (OutUnit& U : P.outs) {
    if (tstep < U.tm || tstep > U.tp || (tstep - U.tm) % U.ts != 0) continue;
    if (U.tp > U.tm) {
      double t = (double)tstep * P.dt;
      fwrite(&t, sizeof(double), 1, U.tfp);
    }
    std::vector<double> buf;
    buf.reserve((size_t)U.nx[0] * U.nx[1] * U.nx[2]);
    if (U.location == -1) {
      for (int a = U.xm[0]; a <= U.xp[0]; a += U.xs[0])
        for (int b = U.xm[1]; b <= U.xp[1]; b += U.xs[1])
          for (int c = U.xm[2]; c <= U.xp[2]; c += U.xs[2]) buf.push_back(P.f[(size_t)U.field * P.N + P.idx(a, b, c)]);
    } else {
      const Iface& I = P.ifaces[U.location];
      const std::vector<double>& arr = *iface_array(P, I, U.field_name);
      int d = I.dir;
      int ax0 = (d == 0) ? 1 : 0, ax1 = (d == 2) ? 1 : 2;
      size_t off = (size_t)U.comp * I.n[0] * I.n[1];
      for (int a = U.xm[0]; a <= U.xp[0]; a += U.xs[0])
        for (int b = U.xm[1]; b <= U.xp[1]; b += U.xs[1])
          for (int c = U.xm[2]; c <= U.xp[2]; c += U.xs[2]) {
            int cc[3] = {a, b, c};
            size_t q = off + (size_t)(cc[ax0] - I.xm_loc[ax0]) * I.n[1] + (cc[ax1] - I.xm_loc[ax1]);
            buf.push_back(arr[q]);
          }
    }
    fwrite(buf.data(), sizeof(double), buf.size(), U.fp);
  }
}

static void close_outputs(Problem& P) {
  for (OutUnit& U : P.outs) {
    if (U.fp) fclose(U.fp);
    if (U.tfp) fclose(U.tfp);
    U.fp = U.tfp = nullptr;
  }
}

static void read_fronts(Problem& P, InText& in) {  // frontlist.cpp:20-80
  in.rewind();
  if (!in.scan("[fdfault.frontlist]")) throw Err("Error reading frontlist from input file");
  int has = in.integer();
  if (!has) return;
  std::string field = in.tok();
  double value = in.dbl();
  if (!(value > 0.) || (field != "V" && field != "U")) throw Err("invalid front parameters");
  for (int i = 0; i < P.nifaces; i++) {
    if (!P.ifaces[i].is_friction) continue;
    Front F;
    F.iface = i;
    F.field = (field == "V") ? 0 : 1;
    F.value = value;
    F.tvals.assign((size_t)P.ifaces[i].n[0] * P.ifaces[i].n[1], -1.);
    P.fronts.push_back(F);
  }
}

static void set_front(Problem& P, double t) {  // front.cpp:103-125
  for (Front& F : P.fronts) {
    const Iface& I = P.ifaces[F.iface];
    for (size_t i = 0; i < F.tvals.size(); i++) {
      if (F.tvals[i] < 0. && F.field == 0 && I.v[i] >= F.value) F.tvals[i] = t;
    }
  }
}

static void write_fronts(const Problem& P) {  // front.cpp:127-359
  const char endian = endian_char();
  for (const Front& F : P.fronts) {
    const Iface& I = P.ifaces[F.iface];
    std::string pre = P.datadir + P.name + "_front_" + std::to_string(F.iface);
    FILE* fp = fopen((pre + "_t.dat").c_str(), "wb");
    if (!fp) throw Err("Error opening file in front.cpp");
    fwrite(F.tvals.data(), sizeof(double), F.tvals.size(), fp);
    fclose(fp);
    const char* xyz[3] = {"x", "y", "z"};
    for (int i = 0; i < P.ndim; i++) {
      FILE* xf = fopen((pre + "_" + xyz[i] + ".dat").c_str(), "wb");
      if (!xf) throw Err("Error opening file in outputunit.cpp");
      for (int a = I.xm_loc[0]; a <= I.xp_loc[0]; a++)
        for (int b = I.xm_loc[1]; b <= I.xp_loc[1]; b++)
          for (int c = I.xm_loc[2]; c <= I.xp_loc[2]; c++) {
            double v = P.x[(size_t)i * P.N + P.idx(a, b, c)];
            fwrite(&v, sizeof(double), 1, xf);
          }
      fclose(xf);
    }
    FILE* m = fopen((pre + ".m").c_str(), "w");
    if (!m) throw Err("Error writing parameters to matlab file");
    fprintf(m, "endian = '%c';\n", endian == '<' ? 'l' : (endian == '>' ? 'b' : 'n'));
    fprintf(m, "nx = %d;\nny = %d;\n", I.n[0], I.n[1]);
    fclose(m);
    FILE* py = fopen((pre + ".py").c_str(), "w");
    if (!py) throw Err("Error writing parameters to python file");
    fprintf(py, "endian = '%c'\n", endian);
    fprintf(py, "nx = %d\nny = %d\n", I.n[0], I.n[1]);
    fclose(py);
  }
}

// problem::problem, problem.cpp:16-84 ; outputs optional (tests step without files)
static Problem* open_problem(const std::string& path, bool with_outputs, bool quiet) {
  Problem* P = new Problem();
  try {
    InText in(path);
    P->inpath = path;
    read_problem(*P, in);
    read_domain(*P, in);
    read_fields(*P, in);
    read_blocks(*P, in);
    read_interfaces(*P, in);
    for (Iface& I : P->ifaces) apply_interface(*P, I, 0., 0., true);
    set_time_step(*P, quiet);
    if (with_outputs) open_outputs(*P, in);
    read_fronts(*P, in);
    set_stress(*P, 1.);
    if (with_outputs) write_outputs(*P, 0);
    set_stress(*P, -1.);
    set_front(*P, 0.);
  } catch (...) {
    close_outputs(*P);
    delete P;
    throw;
  }
  return P;
}

// one full time step, problem.cpp:151-170
static void time_step(Problem& P, bool outputs) {
  int i = P.step;
  for (int st = 0; st < P.rk.nstages; st++) do_rk_stage(P, st, (double)i * P.dt);
  set_stress(P, 1.);
  if (outputs) write_outputs(P, i + 1);
  set_stress(P, -1.);
  set_front(P, (double)(i + 1) * P.dt);
  P.step++;
}

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_open(const char* inpath) {
  try {
    return open_problem(inpath, false, true);
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_close(void* h) {
  if (!h) return;
  Problem* P = (Problem*)h;
  close_outputs(*P);
  delete P;
}

int orc_info(void* h, int* ints, double* dbls) {
  Problem* P = (Problem*)h;
  ints[0] = P->nt; ints[1] = P->rk.nstages; ints[2] = P->ndim; ints[3] = P->mode;
  ints[4] = P->nx[0]; ints[5] = P->nx[1]; ints[6] = P->nx[2];
  ints[7] = P->nf; ints[8] = P->nf + P->nfp; ints[9] = P->nifaces; ints[10] = P->sbporder;
  ints[11] = P->plastic ? 1 : 0; ints[12] = (int)P->blocks.size();
  dbls[0] = P->dt; dbls[1] = P->ttot; dbls[2] = P->cfl;
  return 0;
}

int orc_step(void* h, int nsteps, int write_outputs) {
  Problem* P = (Problem*)h;
  try {
    for (int i = 0; i < nsteps; i++) time_step(*P, write_outputs != 0);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int orc_stage(void* h, int istep, int stage) {
  Problem* P = (Problem*)h;
  try {
    do_rk_stage(*P, stage, (double)istep * P->dt);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

static std::vector<double>* arr_sel(Problem* P, int which) {
  switch (which) {
    case 0: return &P->f;
    case 1: return &P->df;
    case 2: return &P->x;
    case 3: return &P->metric;
    case 4: return &P->jac;
  }
  return nullptr;
}

int orc_get_array(void* h, int which, double* out, long n) {
  std::vector<double>* a = arr_sel((Problem*)h, which);
  if (!a || (size_t)n != a->size()) return -1;
  memcpy(out, a->data(), sizeof(double) * n);
  return 0;
}

int orc_set_array(void* h, int which, const double* in, long n) {
  std::vector<double>* a = arr_sel((Problem*)h, which);
  if (!a || (size_t)n != a->size()) return -1;
  memcpy(a->data(), in, sizeof(double) * n);
  return 0;
}

int orc_iface_shape(void* h, int iface, int* shape) {
  Problem* P = (Problem*)h;
  if (iface < 0 || iface >= P->nifaces) return -1;
  const Iface& I = P->ifaces[iface];
  shape[0] = I.n[0]; shape[1] = I.n[1]; shape[2] = I.dir; shape[3] = I.type;
  return 0;
}

long orc_iface_get(void* h, int iface, int which, double* out, long n) {
  Problem* P = (Problem*)h;
  if (iface < 0 || iface >= P->nifaces) return -1;
  Iface& I = P->ifaces[iface];
  std::vector<double>* a = nullptr;
  switch (which) {
    case 0: a = &I.u; break;
    case 1: a = &I.v; break;
    case 2: a = &I.s; break;
    case 3: a = &I.sn; break;
    case 4: a = &I.ux; break;
    case 5: a = &I.vx; break;
    case 6: a = &I.sx; break;
    case 7: a = &I.state; break;
    case 8: a = &I.du; break;
    case 9: a = &I.dux; break;
    case 10: a = &I.dstate; break;
  }
  if (!a) return -1;
  if (out) {
    if ((size_t)n < a->size()) return -2;
    memcpy(out, a->data(), sizeof(double) * a->size());
  }
  return (long)a->size();
}

long orc_front_get(void* h, int iface, double* out, long n) {
  Problem* P = (Problem*)h;
  for (const Front& F : P->fronts)
    if (F.iface == iface) {
      if (out) {
        if ((size_t)n < F.tvals.size()) return -2;
        memcpy(out, F.tvals.data(), sizeof(double) * F.tvals.size());
      }
      return (long)F.tvals.size();
    }
  return -1;
}

int orc_finish(void* h) {
  Problem* P = (Problem*)h;
  try {
    close_outputs(*P);
    write_fronts(*P);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int orc_run(const char* inpath, int quiet) {
  Problem* P = nullptr;
  try {
    P = open_problem(inpath, true, quiet != 0);
    for (int i = 0; i < P->nt; i++) {
      time_step(*P, true);
      if (!quiet && (i + 1) % P->ninfo == 0) printf("Timestep %d of %d\n", i + 1, P->nt);
    }
    close_outputs(*P);
    write_fronts(*P);
  } catch (std::exception& e) {
    g_err = e.what();
    if (!quiet) fprintf(stderr, "%s\n", e.what());
    if (P) { close_outputs(*P); delete P; }
    return -1;
  }
  delete P;
  return 0;
}

int orc_fd_coeffs(int order, double* fdc, double* diss, double* h0) {
  FD fd;
  try { fd_init(fd, order); } catch (std::exception& e) { g_err = e.what(); return -1; }
  memcpy(fdc, fd.c, sizeof(fd.c));
  memcpy(diss, fd.d, sizeof(fd.d));
  *h0 = fd.h0;
  return 0;
}

int orc_rk_coeffs(int order, int* nstages, double* A, double* B, double* C) {
  RK rk;
  try { rk_init(rk, order); } catch (std::exception& e) { g_err = e.what(); return -1; }
  *nstages = rk.nstages;
  memcpy(A, rk.A, sizeof(rk.A));
  memcpy(B, rk.B, sizeof(rk.B));
  memcpy(C, rk.C, sizeof(rk.C));
  return 0;
}

double orc_stz_func(double mu, const double* p) { return stz_func(mu, p); }

double orc_stz_mu(const double* p, int* status) {
  *status = 0;
  try {
    double phi = p[0], snc = p[2];
    if (p[7] >= -phi / snc) return -phi / snc;
    return solve_newton(p[7], -phi / snc, p);
  } catch (std::exception& e) {
    g_err = e.what();
    *status = -1;
    return NAN;
  }
}

void orc_plastic_flow(double dt, const double* in, double k, double g, double mu, double c, double beta, double eta,
                      double* out) {
  PS si;
  memcpy(&si, in, sizeof(PS));
  PMat m{mu, c, beta, eta};
  PS so = plastic_flow(dt, si, k, g, m);
  memcpy(out, &so, sizeof(PS));
}

double orc_stz_vpl(double mu, const double* p) { return calc_vpl(mu, p); }

double orc_stz_dstatedt(double V, double S, double chi, double c0, double R, double beta, double chiw, double v1,
                        int* status) {
  *status = 0;
  try {
    return stz_dstatedt(V, S, chi, c0, R, beta, chiw, v1);
  } catch (const Err&) {
    *status = -1;
    return 0.;
  }
}

void orc_sw_solve(double phi, double eta, double snc, double c0, double mu, double* vs) {
  sw_solve_point(phi, eta, snc, c0, mu, vs[0], vs[1]);
}

double orc_sw_mu(double u, double t, double dc, double mus, double mud, double trup, double tc) {
  return sw_mu_value(u, t, dc, mus, mud, trup, tc);
}

void orc_solve_locked(double* io, double z1, double z2) {
  IfChar c{io[2], io[3], io[0], io[1]};
  IfChar o = solve_locked(c, z1, z2);
  io[0] = o.s1; io[1] = o.s2; io[2] = o.v1; io[3] = o.v2;
}

void orc_rotate(const double* b9, const double* nn, const double* t1, const double* t2, int to_nt, double* out9) {
  BF b;
  memcpy(&b, b9, sizeof(BF));
  BF o = to_nt ? rotate_xy_nt(b, nn, t1, t2) : rotate_nt_xy(b, nn, t1, t2);
  memcpy(out9, &o, sizeof(BF));
}

// ---- proposal network (test infrastructure for the GPU training kernels) --
// Layer by layer in double precision: h1 = tanh(W1 x + b1),
// h2 = tanh(W2 h1 + b2), y = W3 . h2 + b3; the gradient is plain reverse-mode
// differentiation of the mean squared error over the batch.
static const int NI = 4, NH = 64;
static const int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH,
                 NWTS = OB3 + 1;

static double mlp_one(const double* P, const double* x, double* h1, double* h2) {
  for (int j = 0; j < NH; j++) {
    double a = P[OB1 + j];
    for (int i = 0; i < NI; i++) a += P[OW1 + j * NI + i] * x[i];
    h1[j] = tanh(a);
  }
  double y = P[OB3];
  for (int j = 0; j < NH; j++) {
    double a = P[OB2 + j];
    for (int k = 0; k < NH; k++) a += P[OW2 + j * NH + k] * h1[k];
    h2[j] = tanh(a);
    y += P[OW3 + j] * h2[j];
  }
  return y;
}

void orc_mlp_forward(const double* P, const double* X, long n, double* Y) {
  double h1[NH], h2[NH];
  for (long b = 0; b < n; b++) Y[b] = mlp_one(P, X + b * NI, h1, h2);
}

double orc_mlp_grad(const double* P, const double* X, const double* T, long n, double* G) {
  for (int q = 0; q < NWTS; q++) G[q] = 0.;
  double loss = 0.;
  double h1[NH], h2[NH], d2[NH];
  for (long b = 0; b < n; b++) {
    const double* x = X + b * NI;
    const double r = mlp_one(P, x, h1, h2) - T[b];
    loss += 0.5 * r * r;
    const double dy = r / (double)n;
    G[OB3] += dy;
    for (int j = 0; j < NH; j++) {
      G[OW3 + j] += dy * h2[j];
      d2[j] = dy * P[OW3 + j] * (1. - h2[j] * h2[j]);
      G[OB2 + j] += d2[j];
      for (int k = 0; k < NH; k++) G[OW2 + j * NH + k] += d2[j] * h1[k];
    }
    for (int k = 0; k < NH; k++) {
      double g = 0.;
      for (int j = 0; j < NH; j++) g += d2[j] * P[OW2 + j * NH + k];
      const double d1 = g * (1. - h1[k] * h1[k]);
      G[OB1 + k] += d1;
      for (int i = 0; i < NI; i++) G[OW1 + k * NI + i] += d1 * x[i];
    }
  }
  return loss / (double)n;
}

}  // extern "C"
