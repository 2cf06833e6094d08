// oracle.cpp -- CPU FP64 restatement of DIABLO 3's channel RK3/CN substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg call it to check
// (or time beside) the HIP path.  Nothing in the product links or calls it.
//
// It restates, loop by loop and on the reference's own aliased memory layout
// (one buffer per field, physical real view (0:Nx+1,0:Nz+1,0:Ny+1) and Fourier
// complex view (0:Nx/2,0:Nz+1,0:Ny+1) of the same memory, fft.f90:72-97), the
// single-y-rank (NprocY = NprocZ = 1) behaviour of:
//   for/channel.f90   rk_chan_1 (10-917), rem_div_chan (946-1023), the real
//                     and complex Thomas sweeps (1335-1540), the BC row
//                     routines (1545-2015)
//   for/fft.f90       fft_xz_to_fourier / fft_xz_to_physical (247-382),
//                     wavenumbers (213-230) with pi = 4*atan(1.0) in single
//                     precision (152)
//   for/boundary.f90  apply_BC_vel_{lower,upper}_post (299-498)
//   for/domain.f90    metrics (230-254, 501-527) and jstart/jend (284-303)
//   for/flow.f90      RK3 coefficients (78-86)
//   for/les.f90       les_chan / les_chan_th (in oracle_les.cpp)
// FFTW (not vendored in the reference) is replaced by a textbook radix-2
// Cooley-Tukey FFT (a recursive mixed-radix Cooley-Tukey for lengths that
// are not powers of two, each prime factor by its plain DFT) with FFTW's
// conventions: unnormalised, forward sign -1, r2c keeps k=0..N/2, c2r
// ignores Im of the k=0 and k=N/2 inputs.
//
// Parity status per function (what pins it; see tests/test_oracle_*.py):
//   fft_xz_to_fourier/physical  numpy.fft closed forms + single-mode KATs
//   thomas real/complex         dense numpy.linalg.solve on the same rows
//   rem_div_chan                discrete divergence == 0 invariant
//   rk_chan_1 (DNS)             Wray-RK3 single-mode decay closed form,
//                               Poiseuille fixed point, round trips
#include "../include/diablo_gpu.h"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oracle_internal.h"

namespace orc {

// ---------------------------------------------------------------- FFT ----
void CFFT::init(int n_) {
  n = n_;
  logn = 0;
  while ((1 << logn) < n) ++logn;
  pow2 = (1 << logn) == n;
  if (!pow2) {
    w.resize(n);
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      w[k] = cplx(std::cos(a), std::sin(a));
    }
    return;
  }
  w.resize(n / 2 > 0 ? n / 2 : 1);
  for (int k = 0; k < n / 2; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    w[k] = cplx(std::cos(a), std::sin(a));
  }
  rev.resize(n);
  for (int i = 0; i < n; ++i) {
    int r = 0;
    for (int b = 0; b < logn; ++b)
      if (i & (1 << b)) r |= 1 << (logn - 1 - b);
    rev[i] = r;
  }
}

// Decimation in time over the smallest prime factor p of m (m = n / stride):
// X[k + s m/p] = sum_r w_m^{r (k + s m/p)} DFT_{m/p}(x[r::p])[k].
static void dft_rec(const cplx* x, int stride, cplx* out, int m, int n, const std::vector<cplx>& w,
                    bool inverse) {
  if (m == 1) {
    out[0] = x[0];
    return;
  }
  int p = 2;
  while (m % p) ++p;
  const int q = m / p;
  std::vector<cplx> sub((size_t)m);
  for (int r = 0; r < p; ++r) dft_rec(x + (size_t)r * stride, stride * p, sub.data() + (size_t)r * q, q, n, w, inverse);
  const int unit = n / m;  // w_m^e = w[e * unit mod n]
  for (int k = 0; k < q; ++k)
    for (int s = 0; s < p; ++s) {
      const int kk = k + s * q;
      cplx acc(0.0, 0.0);
      for (int r = 0; r < p; ++r) {
        cplx t = w[(int)(((long long)r * kk * unit) % n)];
        if (inverse) t = std::conj(t);
        acc += sub[(size_t)r * q + k] * t;
      }
      out[kk] = acc;
    }
}

void CFFT::exec_general(cplx* a, bool inverse) const {
  std::vector<cplx> x(a, a + n);
  dft_rec(x.data(), 1, a, n, n, w, inverse);
}

void CFFT::exec(cplx* a, bool inverse) const {
  if (!pow2) {
    exec_general(a, inverse);
    return;
  }
  for (int i = 0; i < n; ++i)
    if (i < rev[i]) std::swap(a[i], a[rev[i]]);
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int j = 0; j < half; ++j) {
        cplx t = w[j * step];
        if (inverse) t = std::conj(t);
        cplx u = a[i + j], v = a[i + j + half] * t;
        a[i + j] = u + v;
        a[i + j + half] = u - v;
      }
    }
  }
}

// r2c of two real rows at once: Z = FFT(a + i b).
void real_pair_forward(const CFFT& f, const double* a, const double* b, cplx* A, cplx* B,
                       cplx* work) {
  int n = f.n;
  for (int x = 0; x < n; ++x) work[x] = cplx(a[x], b ? b[x] : 0.0);
  f.exec(work, false);
  for (int k = 0; k <= n / 2; ++k) {
    cplx zk = work[k % n], znk = std::conj(work[(n - k) % n]);
    A[k] = 0.5 * (zk + znk);
    if (B) B[k] = cplx(0.0, -0.5) * (zk - znk);
  }
  A[0] = cplx(A[0].real(), 0.0);
  if (B) B[0] = cplx(B[0].real(), 0.0);
}

// c2r of two Hermitian half-spectra at once (FFTW c2r semantics: Im of the
// k=0 and k=N/2 inputs is ignored).
void real_pair_inverse(const CFFT& f, const cplx* A, const cplx* B, double* a, double* b,
                       cplx* work) {
  int n = f.n;
  auto hA = [&](int k) -> cplx {
    if (k == 0 || k == n / 2) return cplx(A[k].real(), 0.0);
    if (k < n / 2) return A[k];
    return std::conj(A[n - k]);
  };
  auto hB = [&](int k) -> cplx {
    if (!B) return cplx(0.0, 0.0);
    if (k == 0 || k == n / 2) return cplx(B[k].real(), 0.0);
    if (k < n / 2) return B[k];
    return std::conj(B[n - k]);
  };
  for (int k = 0; k < n; ++k) {
    cplx bb = hB(k);
    work[k] = hA(k) + cplx(-bb.imag(), bb.real());
  }
  f.exec(work, true);
  for (int x = 0; x < n; ++x) {
    a[x] = work[x].real();
    if (b) b[x] = work[x].imag();
  }
}

// ------------------------------------------------------------ Solver ----
Solver::Solver(const dgpu_params& prm, const double* gy_in) : p(prm) {
  Nx = p.nx;
  Ny = p.ny;
  Nz = p.nz;
  Nth = p.n_th;
  Nyp = Ny;          // (Ny-1)/NprocY + 1 at NprocY=1 (domain.f90:12)
  Nxp = Nx / 2;      // domain.f90:13
  Nzp = Nz;
  Nkx = Nx / 3;      // domain.f90:18
  Nkz = Nz / 3;
  twoNkz = 2 * Nkz;
  NX2 = Nx + 2;
  NXC = Nx / 2 + 1;
  NZ2 = Nz + 2;
  NY2 = Ny + 2;
  NB = (size_t)NX2 * NZ2 * NY2;
  nu = 1.0 / p.re;

  // grid: gy(1..Nyp+1) from grid.h5 (domain.f90:429-483); gy(0) is never read
  gy.assign(NY2, 0.0);
  gyf.assign(NY2, 0.0);
  dy.assign(NY2, 0.0);
  dyf.assign(NY2, 0.0);
  for (int j = 1; j <= Nyp + 1; ++j) gy[j] = gy_in[j - 1];
  for (int j = 1; j <= Nyp; ++j) gyf[j] = 0.5 * (gy[j] + gy[j + 1]);    // domain.f90:501-503
  gyf[0] = 2.0 * gyf[1] - gyf[2];                                      // :511
  gyf[Nyp + 1] = 2.0 * gyf[Nyp] - gyf[Nyp - 1];                        // :520
  for (int j = 1; j <= Nyp + 1; ++j) dy[j] = gyf[j] - gyf[j - 1];      // :230-232
  for (int j = 1; j <= Nyp; ++j) dyf[j] = gy[j + 1] - gy[j];           // :233-235
  dyf[0] = dyf[1];                                                      // :240
  dyf[Nyp + 1] = dyf[Nyp];                                              // :248

  // wavenumbers (fft.f90:152, 213-230): pi is evaluated in default real
  pi = (double)(4.0f * std::atan(1.0f));
  kx.assign(Nxp + 1, 0.0);
  kx2.assign(Nxp + 1, 0.0);
  for (int i = 0; i < Nxp; ++i) {
    kx[i] = (double)i * (2.0 * pi) / p.lx;
    kx2[i] = kx[i] * kx[i];
  }
  kz.assign(twoNkz + 1, 0.0);
  kz2.assign(twoNkz + 1, 0.0);
  for (int k = 0; k <= Nkz; ++k) kz[k] = (double)k * (2.0 * pi) / p.lz;
  for (int k = 1; k <= Nkz; ++k) kz[twoNkz + 1 - k] = -(double)k * (2.0 * pi) / p.lz;
  for (int k = 0; k <= twoNkz; ++k) kz2[k] = kz[k] * kz[k];

  // j-ranges at one y-rank: the rankY==0 branch of init_chan_mpi (domain.f90:284-303)
  jend = Nyp;
  jstart = (p.u_bc_ymin == 1) ? 1 : 2;
  for (int n = 0; n < Nth; ++n) {
    jend_th[n] = Nyp;
    jstart_th[n] = (p.th_bc_ymin[n] == 1) ? 1 : 2;
  }

  // RK3 (flow.f90:78-86)
  h_bar[0] = p.delta_t * (8.0 / 15.0);
  h_bar[1] = p.delta_t * (2.0 / 15.0);
  h_bar[2] = p.delta_t * (5.0 / 15.0);
  beta_bar[0] = 1.0;
  beta_bar[1] = 25.0 / 8.0;
  beta_bar[2] = 9.0 / 4.0;
  zeta_bar[0] = 0.0;
  zeta_bar[1] = -17.0 / 8.0;
  zeta_bar[2] = -5.0 / 4.0;
  w_bc_ymax_c1_transient = p.w_bc_ymax_c1;  // parameters.f90:270

  for (auto* v : {&u1, &u2, &u3, &pp, &r1, &r2, &r3, &f1, &f2, &f3, &s1}) v->assign(NB, 0.0);
  th.assign(NB * (Nth > 0 ? Nth : 1), 0.0);
  fth.assign(NB * (Nth > 0 ? Nth : 1), 0.0);
  rth.assign(NB * (Nth > 0 ? Nth : 1), 0.0);
  nu_t.assign(NB, 0.0);
  kappa_t.assign(NB * (Nth > 0 ? Nth : 1), 0.0);
  temp_fft.assign((size_t)NXC * NZ2 * NY2, cplx(0, 0));
  fx.init(Nx);
  fz.init(Nz);
}

// fft_xz_to_fourier (fft.f90:247-315), in place on one field buffer.
void Solver::fft_xz_to_fourier(double* v) {
  cplx* vv = reinterpret_cast<cplx*>(v);
  // r2c in x over all j=0..Nyp+1, k=0..Nzp-1, then /Nx and zero i in (Nkx, Nx/2]
#pragma omp parallel
  {
    std::vector<cplx> work(Nx), A(Nx / 2 + 1), B(Nx / 2 + 1);
    // rows in pairs; with odd Nzp the last row goes alone
    const int npair = (Nzp + 1) / 2;
#pragma omp for schedule(static)
    for (int jk = 0; jk < NY2 * npair; ++jk) {
      int j = jk / npair, k = 2 * (jk % npair);
      const bool two = k + 1 < Nzp;
      real_pair_forward(fx, &R(v, 0, k, j), two ? &R(v, 0, k + 1, j) : nullptr, A.data(), two ? B.data() : nullptr,
                        work.data());
      for (int kk = 0; kk < (two ? 2 : 1); ++kk) {
        const cplx* src = kk ? B.data() : A.data();
        for (int i = 0; i <= Nkx; ++i) TF(i, k + kk, j) = src[i] / (double)Nx;
        for (int i = Nkx + 1; i <= Nx / 2; ++i) TF(i, k + kk, j) = cplx(0.0, 0.0);
      }
    }
  }
  // MPI_Alltoall at NprocZ=1: copy temp_fft(0:Nxp-1, 0:Nz-1, :) into vv (domain.f90:151-195)
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY2; ++j)
    for (int k = 0; k < Nz; ++k)
      for (int i = 0; i < Nxp; ++i) vv[ci(i, k, j)] = TF(i, k, j);
  // c2c forward in z for i=0..Nxp-1, then /Nz and pack (fft.f90:298-313)
#pragma omp parallel
  {
    std::vector<cplx> col(Nz);
#pragma omp for schedule(static)
    for (int ji = 0; ji < NY2 * Nxp; ++ji) {
      int j = ji / Nxp, i = ji % Nxp;
      for (int k = 0; k < Nz; ++k) col[k] = vv[ci(i, k, j)];
      fz.exec(col.data(), false);
      for (int k = 0; k < Nz; ++k) vv[ci(i, k, j)] = col[k];
      for (int k = 0; k <= Nkz; ++k) vv[ci(i, k, j)] = vv[ci(i, k, j)] / (double)Nz;
      for (int k = 1; k <= Nkz; ++k)
        vv[ci(i, Nkz + k, j)] = vv[ci(i, Nz - 1 + k - Nkz, j)] / (double)Nz;
    }
  }
}

// fft_xz_to_physical (fft.f90:319-382), in place; destroys the Fourier data.
void Solver::fft_xz_to_physical(double* v) {
  cplx* vv = reinterpret_cast<cplx*>(v);
#pragma omp parallel
  {
    std::vector<cplx> col(Nz);
#pragma omp for schedule(static)
    for (int ji = 0; ji < NY2 * Nxp; ++ji) {
      int j = ji / Nxp, i = ji % Nxp;
      for (int k = Nkz; k >= 1; --k) vv[ci(i, Nz - 1 + k - Nkz, j)] = vv[ci(i, Nkz + k, j)];
      for (int k = Nkz + 1; k <= Nz - Nkz - 1; ++k) vv[ci(i, k, j)] = cplx(0.0, 0.0);
      for (int k = 0; k < Nz; ++k) col[k] = vv[ci(i, k, j)];
      fz.exec(col.data(), true);
      for (int k = 0; k < Nz; ++k) vv[ci(i, k, j)] = col[k];
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < NY2; ++j)
    for (int k = 0; k < Nz; ++k)
      for (int i = 0; i < Nxp; ++i) TF(i, k, j) = vv[ci(i, k, j)];
#pragma omp parallel
  {
    std::vector<cplx> work(Nx);
    const int npair = (Nzp + 1) / 2;
#pragma omp for schedule(static)
    for (int jk = 0; jk < NY2 * npair; ++jk) {
      int j = jk / npair, k = 2 * (jk % npair);
      const bool two = k + 1 < Nzp;
      for (int kk = 0; kk < (two ? 2 : 1); ++kk)
        for (int i = Nkx + 1; i <= Nx / 2; ++i) TF(i, k + kk, j) = cplx(0.0, 0.0);
      real_pair_inverse(fx, &TF(0, k, j), two ? &TF(0, k + 1, j) : nullptr, &R(v, 0, k, j),
                        two ? &R(v, 0, k + 1, j) : nullptr, work.data());
    }
  }
}

// thomas_forward_real_mpi + thomas_backward_real_mpi at one y-rank
// (channel.f90:1335-1405, 1460-1496); rows 0..Nyp+1, lanes 0..inx-1.
void thomas_real(double* a, double* b, double* c, double* g, int Nyp, int inx) {
  for (int j = 1; j <= Nyp; ++j)
    for (int i = 0; i < inx; ++i) {
      size_t o = i + (size_t)inx * j, m = o - inx;
      a[o] = -a[o] / b[m];
      b[o] = b[o] + a[o] * c[m];
      g[o] = g[o] + a[o] * g[m];
    }
  {
    int j = Nyp + 1;
    for (int i = 0; i < inx; ++i) {
      size_t o = i + (size_t)inx * j, m = o - inx;
      a[o] = -a[o] / b[m];
      b[o] = b[o] + a[o] * c[m];
      g[o] = g[o] + a[o] * g[m];
    }
  }
  for (int i = 0; i < inx; ++i) {
    size_t t = i + (size_t)inx * (Nyp + 1);
    g[t] = g[t] / b[t];
    for (int j = Nyp; j >= 0; --j) {
      size_t o = i + (size_t)inx * j;
      g[o] = (g[o] - c[o] * g[o + inx]) / b[o];
    }
  }
}

// thomas_forward_complex_mpi + thomas_backward_complex_mpi at one y-rank
// (channel.f90:1408-1457, 1499-1540): forward j=2..Nyp, backward Nyp..1.
void thomas_complex(double* a, double* b, double* c, cplx* g, int Nyp, int inx) {
  for (int i = 0; i < inx; ++i) {
    for (int j = 2; j <= Nyp; ++j) {
      size_t o = i + (size_t)inx * j, m = o - inx;
      a[o] = -a[o] / b[m];
      b[o] = b[o] + a[o] * c[m];
      g[o] = g[o] + a[o] * g[m];
    }
    size_t t = i + (size_t)inx * Nyp;
    g[t] = g[t] / b[t];
    for (int j = Nyp - 1; j >= 1; --j) {
      size_t o = i + (size_t)inx * j;
      g[o] = (g[o] - c[o] * g[o + inx]) / b[o];
    }
  }
}

// rem_div_chan (channel.f90:946-1023) with apply_BC_rem_div_mpi (:1974-2015).
void Solver::rem_div_chan() {
  cplx* cu1 = C(u1);
  cplx* cu2 = C(u2);
  cplx* cu3 = C(u3);
  cplx* cr1 = C(r1);
  int inx = Nxp;  // lanes 0..Nxp-1 carry data (the reference also sweeps an uninitialised lane Nxp)
#pragma omp parallel
  {
    std::vector<double> ml((size_t)inx * NY2), md((size_t)inx * NY2), mu((size_t)inx * NY2);
    std::vector<cplx> vc((size_t)inx * NY2);
#pragma omp for schedule(dynamic)
    for (int k = 0; k <= twoNkz; ++k) {
      for (int j = 0; j < NY2; ++j)
        for (int i = 0; i < inx; ++i) {
          size_t o = i + (size_t)inx * j;
          ml[o] = 0.0;
          md[o] = 1.0;
          mu[o] = 0.0;
          vc[o] = cplx(0.0, 0.0);
        }
      for (int j = 1; j <= Nyp; ++j)
        for (int i = 0; i < inx; ++i) {
          size_t o = i + (size_t)inx * j;
          ml[o] = 1.0 / (dy[j] * dyf[j]);
          md[o] = -kx2[i] - kz2[k] - 1.0 / (dy[j + 1] * dyf[j]) - 1.0 / (dy[j] * dyf[j]);
          mu[o] = 1.0 / (dy[j + 1] * dyf[j]);
        }
      for (int j = 1; j <= Nyp; ++j)
        for (int i = 0; i < inx; ++i) {
          size_t o = i + (size_t)inx * j;
          vc[o] = ik(kx[i], cu1[ci(i, k, j)]) + (cu2[ci(i, k, j + 1)] - cu2[ci(i, k, j)]) / dyf[j] +
                  ik(kz[k], cu3[ci(i, k, j)]);
        }
      // apply_BC_rem_div_mpi (channel.f90:1974-2015)
      for (int i = 0; i < inx; ++i) {
        size_t o = i + (size_t)inx * 1;
        ml[o] = 0.0;
        md[o] = 1.0;
        mu[o] = (k == 0 && i == 0) ? 0.0 : -1.0;
        vc[o] = cplx(0.0, 0.0);
        size_t t = i + (size_t)inx * Nyp;
        ml[t] = 1.0;
        md[t] = -1.0;
        mu[t] = 0.0;
        vc[t] = cplx(0.0, 0.0);
      }
      thomas_complex(ml.data(), md.data(), mu.data(), vc.data(), Nyp, inx);
      for (int j = 1; j <= Nyp; ++j)
        for (int i = 0; i < inx; ++i) cr1[ci(i, k, j)] = vc[i + (size_t)inx * j];
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu1[ci(i, k, j)] -= ik(kx[i], cr1[ci(i, k, j)]);
        cu3[ci(i, k, j)] -= ik(kz[k], cr1[ci(i, k, j)]);
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i)
        cu2[ci(i, k, j)] -= (cr1[ci(i, k, j)] - cr1[ci(i, k, j - 1)]) / dy[j];
}

// apply_BC_vel_lower_post / _upper_post (boundary.f90:299-498)
void Solver::apply_bc_vel_post(double time) {
  cplx* cu1 = C(u1);
  cplx* cu2 = C(u2);
  cplx* cu3 = C(u3);
  const cplx Z(0.0, 0.0);
  // lower wall
  if (p.u_bc_ymin == 0) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cu1[ci(i, k, 0)] = cu1[ci(i, k, 1)] = Z;
    cu1[ci(0, 0, 1)] = p.u_bc_ymin_c1;
    cu1[ci(0, 0, 0)] = p.u_bc_ymin_c1;
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu1[ci(i, k, 1)] = cu1[ci(i, k, 2)];
        cu1[ci(i, k, 0)] = cu1[ci(i, k, 1)];
      }
    cu1[ci(0, 0, 1)] = cu1[ci(0, 0, 2)] - dy[2] * p.u_bc_ymin_c1;
    cu1[ci(0, 0, 0)] = cu1[ci(0, 0, 1)] - dy[1] * p.u_bc_ymin_c1;
  }
  if (p.w_bc_ymin == 0) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cu3[ci(i, k, 0)] = cu3[ci(i, k, 1)] = Z;
    cu3[ci(0, 0, 1)] = p.w_bc_ymin_c1;
    cu3[ci(0, 0, 0)] = p.w_bc_ymin_c1;
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu3[ci(i, k, 1)] = cu3[ci(i, k, 2)];
        cu3[ci(i, k, 0)] = cu3[ci(i, k, 1)];
      }
    cu3[ci(0, 0, 1)] = cu3[ci(0, 0, 2)] - dy[2] * p.w_bc_ymin_c1;
    cu3[ci(0, 0, 0)] = cu3[ci(0, 0, 1)] - dy[1] * p.w_bc_ymin_c1;
  }
  if (p.v_bc_ymin == 0) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu2[ci(i, k, 1)] = 2.0 * p.v_bc_ymin_c1 - cu2[ci(i, k, 2)];
        cu2[ci(i, k, 0)] = cu2[ci(i, k, 1)];
      }
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu2[ci(i, k, 1)] = cu2[ci(i, k, 2)];
        cu2[ci(i, k, 0)] = cu2[ci(i, k, 1)];
      }
    cu2[ci(0, 0, 1)] = cu2[ci(0, 0, 2)] - dyf[1] * p.v_bc_ymin_c1;
    cu2[ci(0, 0, 0)] = cu2[ci(0, 0, 1)] - dyf[1] * p.v_bc_ymin_c1;
  }
  // upper wall
  if (p.u_bc_ymax == 0) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cu1[ci(i, k, Nyp)] = cu1[ci(i, k, Nyp + 1)] = Z;
    cu1[ci(0, 0, Nyp)] = p.u_bc_ymax_c1;
    cu1[ci(0, 0, Nyp + 1)] = p.u_bc_ymax_c1;
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu1[ci(i, k, Nyp)] = cu1[ci(i, k, Nyp - 1)];
        cu1[ci(i, k, Nyp + 1)] = cu1[ci(i, k, Nyp)];
      }
    cu1[ci(0, 0, Nyp)] = cu1[ci(0, 0, Nyp - 1)] + dy[Nyp] * p.u_bc_ymax_c1;
    cu1[ci(0, 0, Nyp + 1)] = cu1[ci(0, 0, Nyp)] + dy[Nyp] * p.u_bc_ymax_c1;
  }
  if (p.w_bc_ymax == 0) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cu3[ci(i, k, Nyp)] = cu3[ci(i, k, Nyp + 1)] = Z;
    cu3[ci(0, 0, Nyp)] = p.w_bc_ymax_c1;
    cu3[ci(0, 0, Nyp + 1)] = p.w_bc_ymax_c1;
    cu3[ci(0, 0, Nyp + 1)] = Z;  // "Ghost cell not used" (boundary.f90:442)
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu3[ci(i, k, Nyp)] = cu3[ci(i, k, Nyp - 1)];
        cu3[ci(i, k, Nyp + 1)] = cu3[ci(i, k, Nyp)];
      }
    if (p.f_type == 4)
      w_bc_ymax_c1_transient =
          p.w_bc_ymax_c1 +
          p.amp_omega0 * std::sin(p.omega0 * (p.ro_inv / p.delta * time - p.force_start));
    cu3[ci(0, 0, Nyp)] = cu3[ci(0, 0, Nyp - 1)] + dy[Nyp] * w_bc_ymax_c1_transient;
    cu3[ci(0, 0, Nyp + 1)] = cu3[ci(0, 0, Nyp)] + dy[Nyp] * w_bc_ymax_c1_transient;
  }
  if (p.v_bc_ymax == 0) {
    // only the (0,0) mode is written inside the (i,k) loop (boundary.f90:476-480)
    cu2[ci(0, 0, Nyp + 1)] = 2.0 * p.v_bc_ymax_c1 - cu2[ci(0, 0, Nyp)];
  } else {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cu2[ci(i, k, Nyp)] = cu2[ci(i, k, Nyp - 1)];
        cu2[ci(i, k, Nyp + 1)] = cu2[ci(i, k, Nyp)];
      }
    cu2[ci(0, 0, Nyp + 1)] = cu2[ci(0, 0, Nyp)] + dy[Nyp] * p.v_bc_ymax_c1;
  }
}

// BC rows of the real implicit solves (channel.f90:1545-1876).  kind: 1 = u1,
// 2 = u2, 3 = u3, 4+n = scalar n.  Arrays are (0:Nx-1, 0:Nyp+1).
void Solver::apply_bc_rows(int kind, double* ml, double* md, double* mu, double* vec,
                           double time) {
  int nx = Nx;
  auto row = [&](int j, double l, double d, double u, double v) {
    for (int i = 0; i < nx; ++i) {
      size_t o = i + (size_t)nx * j;
      ml[o] = l;
      md[o] = d;
      mu[o] = u;
      vec[o] = v;
    }
  };
  if (kind == 1 || kind == 3 || kind >= 4) {
    int bcl, bcu;
    double c1l, c1u;
    if (kind == 1) {
      bcl = p.u_bc_ymin; bcu = p.u_bc_ymax; c1l = p.u_bc_ymin_c1; c1u = p.u_bc_ymax_c1;
    } else if (kind == 3) {
      bcl = p.w_bc_ymin; bcu = p.w_bc_ymax; c1l = p.w_bc_ymin_c1; c1u = p.w_bc_ymax_c1;
      if (bcu != 0) {
        if (p.f_type == 4)  // channel.f90:1775
          w_bc_ymax_c1_transient =
              p.w_bc_ymax_c1 +
              p.amp_omega0 * std::sin(p.omega0 * (p.ro_inv / p.delta * time - p.force_start));
        c1u = w_bc_ymax_c1_transient;
      }
    } else {
      int n = kind - 4;
      bcl = p.th_bc_ymin[n]; bcu = p.th_bc_ymax[n];
      c1l = p.th_bc_ymin_c1[n]; c1u = p.th_bc_ymax_c1[n];
    }
    if (bcl == 0) {
      row(0, 0.0, 1.0, 0.0, 0.0);
      row(1, 0.0, 1.0, 0.0, c1l);
    } else {
      row(0, 0.0, 1.0, 0.0, 0.0);
      row(1, 0.0, -1.0, 1.0, dy[2] * c1l);
    }
    if (bcu == 0) {
      row(Nyp + 1, 0.0, 1.0, 0.0, 0.0);
      row(Nyp, 0.0, 1.0, 0.0, c1u);
    } else {
      row(Nyp, -1.0, 1.0, 0.0, dy[Nyp] * c1u);
      row(Nyp + 1, 0.0, 1.0, 0.0, 0.0);
    }
  } else {  // u2 (channel.f90:1636-1708)
    if (p.v_bc_ymin == 0) {
      row(1, 0.0, 1.0, 0.0, p.v_bc_ymin_c1);
      row(2, 0.0, 1.0, 0.0, p.v_bc_ymin_c1);
    } else if (p.v_bc_ymin == 1) {
      row(1, 0.0, -1.0, 1.0, dyf[1] * p.v_bc_ymin_c1);
    }
    row(0, 0.0, 1.0, 0.0, 0.0);
    if (p.v_bc_ymax == 0) {
      row(Nyp + 1, 0.0, 1.0, 0.0, p.v_bc_ymax_c1);
      row(Nyp, 0.0, 1.0, 0.0, p.v_bc_ymax_c1);
    } else if (p.v_bc_ymax == 1) {
      row(Nyp + 1, -1.0, 1.0, 0.0, dyf[Nyp] * p.v_bc_ymax_c1);
    }
  }
}

}  // namespace orc
