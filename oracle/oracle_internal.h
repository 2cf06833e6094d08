// oracle_internal.h -- shared declarations of the CPU oracle (test infrastructure only).
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "../include/diablo_gpu.h"

namespace orc {

typedef std::complex<double> cplx;

struct CFFT {
  int n = 0, logn = 0;
  bool pow2 = true;       // radix-2 path; otherwise mixed-radix recursion
  std::vector<cplx> w;    // exp(-2 pi i k / n): k < n/2 (radix 2) or k < n
  std::vector<int> rev;
  void init(int n_);
  void exec(cplx* a, bool inverse) const;
  void exec_general(cplx* a, bool inverse) const;
};

void real_pair_forward(const CFFT& f, const double* a, const double* b, cplx* A, cplx* B,
                       cplx* work);
void real_pair_inverse(const CFFT& f, const cplx* A, const cplx* B, double* a, double* b,
                       cplx* work);
void thomas_real(double* a, double* b, double* c, double* g, int Nyp, int inx);
void thomas_complex(double* a, double* b, double* c, cplx* g, int Nyp, int inx);

// i*k*c as the reference's cikx(i)*c with cikx = (0,kx) (fft.f90:216)
inline cplx ik(double k, const cplx& c) { return cplx(-k * c.imag(), k * c.real()); }

struct Solver {
  dgpu_params p;
  int Nx, Ny, Nz, Nth, Nyp, Nxp, Nzp, Nkx, Nkz, twoNkz;
  int NX2, NXC, NZ2, NY2;
  size_t NB;  // doubles per field buffer
  double nu, pi;
  std::vector<double> gy, gyf, dy, dyf, kx, kx2, kz, kz2;
  int jstart, jend, jstart_th[DGPU_MAX_TH], jend_th[DGPU_MAX_TH];
  double h_bar[3], beta_bar[3], zeta_bar[3];
  double w_bc_ymax_c1_transient;
  // field buffers (physical and Fourier views of the same memory)
  std::vector<double> u1, u2, u3, pp, r1, r2, r3, f1, f2, f3, s1;
  std::vector<double> th, fth, rth;  // n_th consecutive buffers
  std::vector<double> nu_t, kappa_t;
  std::vector<cplx> temp_fft;
  CFFT fx, fz;
  std::string err;

  Solver(const dgpu_params& prm, const double* gy_in);

  inline size_t ri(int i, int k, int j) const { return i + (size_t)NX2 * (k + (size_t)NZ2 * j); }
  inline size_t ci(int i, int k, int j) const { return i + (size_t)NXC * (k + (size_t)NZ2 * j); }
  inline double& R(double* a, int i, int k, int j) { return a[ri(i, k, j)]; }
  inline double& R(std::vector<double>& a, int i, int k, int j) { return a[ri(i, k, j)]; }
  inline cplx& TF(int i, int k, int j) { return temp_fft[ci(i, k, j)]; }
  inline cplx* C(std::vector<double>& a) { return reinterpret_cast<cplx*>(a.data()); }
  inline double* TH(std::vector<double>& a, int n) { return a.data() + NB * n; }
  inline cplx* CTH(std::vector<double>& a, int n) { return reinterpret_cast<cplx*>(a.data() + NB * n); }

  void fft_xz_to_fourier(double* v);
  void fft_xz_to_physical(double* v);
  void rem_div_chan();
  void apply_bc_vel_post(double time);
  void apply_bc_rows(int kind, double* ml, double* md, double* mu, double* vec, double time);
  void rk_chan_1(int rk_step, double time, int time_step);
  // LES (oracle_les.cpp)
  std::vector<double> les_buf;  // lazily allocated LES work buffers
  bool les_ready = false;
  double* lb(int id);           // LES work buffer id (physical/Fourier views)
  void apply_bc_th_post();
  void apply_bc_les();
  void compute_strain(bool invariant, double beta_sgs);
  void compute_rotation(bool invariant, double beta_sgs);
  void compute_all_gradients();
  int les_chan(double time);
  int les_chan_th(double time);
  // start-up and time-step control (oracle_init.cpp)
  bool var_dt = false;
  dgpu_courant_params cour{};
  void poisson_p_chan(const double* ri);
  double courant_points(const double* a1, const double* a2, const double* a3);
  void courant(const double* a1, const double* a2, const double* a3, double time);
};

}  // namespace orc
