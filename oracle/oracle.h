/*
 * fdfault CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch FP64 C++ restatement of the reference solver's time-stepping
 * core (domain::do_rk_stage, src/domain.cpp:228-305, and everything it calls),
 * of the problem set-up that feeds it, and of the output/front file writers.
 * Each routine cites the reference file:line it follows, and keeps the
 * reference's operation order so that results agree to the last bit where the
 * compiler allows it.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library.  The product (fdfault_amd/) never links it: it is the
 * checker, never the thing measured or shipped.
 *
 * Parity pinning: every function is checked in tests/test_oracle_*.py against
 * closed forms, invariants, brute force or the reference's own example inputs
 * (the reference ships no golden outputs).  Functions without such a check are
 * listed as "parity unpinned" in the repository's design notes; at present
 * none are.
 */
#ifndef FDF_ORACLE_H
#define FDF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* problem handle ------------------------------------------------------- */
void* orc_open(const char* inpath);               /* NULL on error, see orc_last_error */
const char* orc_last_error(void);
void orc_close(void* h);
/* ints:  [0]nt [1]nstages [2]ndim [3]mode [4..6]nx [7]nfields(evolved) [8]nfields total
 *        [9]nifaces [10]sbporder [11]plastic [12]nblocks
 * dbls:  [0]dt [1]ttot [2]cfl */
int orc_info(void* h, int* ints, double* dbls);
/* run nsteps full time steps (all RK stages, plasticity, stress round trip,
 * front update, and outputs if write_outputs) */
int orc_step(void* h, int nsteps, int write_outputs);
/* one RK stage of step istep (no end-of-step work); for kernel-level parity */
int orc_stage(void* h, int istep, int stage);
/* copy arrays: which = 0 f, 1 df, 2 x, 3 metric, 4 jac; n = element count */
int orc_get_array(void* h, int which, double* out, long n);
int orc_set_array(void* h, int which, const double* in, long n);
/* interface arrays: which = 0 U,1 V,2 S,3 Sn,4 Ux,5 Vx,6 Sx,7 state,8 du,9 dux,10 dstate;
 * returns element count (or <0 on error); n0/n1 reported in ints[0..1] */
long orc_iface_get(void* h, int iface, int which, double* out, long n);
int orc_iface_shape(void* h, int iface, int* shape /* n0 n1 dir type */);
long orc_front_get(void* h, int iface, double* out, long n);
/* writes final files (closes outputs, writes fronts) */
int orc_finish(void* h);
/* full run equivalent to `fdfault <in>` : returns 0 or -1 */
int orc_run(const char* inpath, int quiet);

/* unit-level entry points (pinning tests) ---------------------------- */
int orc_fd_coeffs(int order, double* fd /*7x9*/, double* diss /*7x9*/, double* h0);
int orc_rk_coeffs(int order, int* nstages, double* A, double* B, double* C);
/* params: phi eta snc chi v0 f0 a muy ; returns mu ; *status = 0 ok, <0 abort path */
double orc_stz_mu(const double* params8, int* status);
double orc_stz_func(double mu, const double* params8);
/* plastic_flow on (sxx sxy sxz syy syz szz gammap lambda epxx epxy epxz epyy epyz epzz) */
void orc_plastic_flow(double dt, const double* s_in14, double k, double g,
                      double mu, double c, double beta, double eta, double* s_out14);
/* STZ plastic slip rate V(mu) (stz.cpp:436-460), params as orc_stz_mu */
double orc_stz_vpl(double mu, const double* params8);
/* STZ effective-temperature rate (stz.cpp:311-356); *status = -1 when V > v1 */
double orc_stz_dstatedt(double V, double S, double chi, double c0, double R, double beta, double chiw, double v1,
                        int* status);
/* slip-weakening solve for one point (slipweak.cpp:206-251) given the
 * friction coefficient mu and cohesion c0: vs = (V, S) */
void orc_sw_solve(double phi, double eta, double snc, double c0, double mu, double* vs);
/* slip-weakening friction coefficient (slipweak.cpp:253-308) */
double orc_sw_mu(double u, double t, double dc, double mus, double mud, double trup, double tc);
/* solve_locked on one characteristic (interface.cpp:1017-1028): io = s1 s2 v1 v2 */
void orc_solve_locked(double* io4, double z1, double z2);
/* rotate xyz->nt and back (boundary.cpp:505-568); b9 = v1 v2 v3 s11 s12 s13 s22 s23 s33 */
void orc_rotate(const double* b9, const double* nn, const double* t1, const double* t2, int to_nt, double* out9);

/* The learned STZ proposal's network (no counterpart in the reference):
 * 4 -> 64 -> 64 -> 1 with tanh, parameters laid out as
 * W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3 (+ normalisation, unused
 * here).  Forward of n normalised inputs X[n][4] into Y[n]; gradient G[4545]
 * of the batch loss mean((y - t)^2 / 2), returned value = that loss. */
void orc_mlp_forward(const double* P, const double* X, long n, double* Y);
double orc_mlp_grad(const double* P, const double* X, const double* T, long n, double* G);

#ifdef __cplusplus
}
#endif
#endif
