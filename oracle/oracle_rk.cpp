// oracle_rk.cpp -- rk_chan_1 (for/channel.f90:10-917) restated on the CPU, and
// the orc_* C-ABI of the oracle.  TEST INFRASTRUCTURE ONLY (see oracle.cpp).
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "oracle_internal.h"

namespace orc {

void Solver::rk_chan_1(int rk_step, double time, int time_step) {
  const int rs = rk_step - 1;
  const double temp1 = 0.5 * nu * h_bar[rs];     // channel.f90:29-33
  const double temp2 = 0.5 * h_bar[rs];
  const double temp3 = zeta_bar[rs] * h_bar[rs];
  const double temp4 = h_bar[rs];
  const double temp5 = beta_bar[rs] * h_bar[rs];
  const double nvs = p.nu_v_scale;
  cplx *cu1 = C(u1), *cu2 = C(u2), *cu3 = C(u3), *cp = C(pp);
  cplx *cr1 = C(r1), *cr2 = C(r2), *cr3 = C(r3);
  cplx *cf1 = C(f1), *cf2 = C(f2), *cf3 = C(f3), *cs1 = C(s1);
  const int js = jstart, je = jend;

  // ---- explicit spectral RHS (channel.f90:39-253)
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cr1[ci(i, k, j)] = cu1[ci(i, k, j)];
        cr3[ci(i, k, j)] = cu3[ci(i, k, j)];
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cr2[ci(i, k, j)] = cu2[ci(i, k, j)];
  if (rk_step > 1) {
#pragma omp parallel for schedule(static)
    for (int j = js; j <= je; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          cr1[ci(i, k, j)] += temp3 * cf1[ci(i, k, j)];
          cr3[ci(i, k, j)] += temp3 * cf3[ci(i, k, j)];
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) cr2[ci(i, k, j)] += temp3 * cf2[ci(i, k, j)];
  }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i)
        cs1[ci(i, k, j)] = (cp[ci(i, k, j)] - cp[ci(i, k, j - 1)]) / dy[j];
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cr1[ci(i, k, j)] -= temp4 * ik(kx[i], cp[ci(i, k, j)]);
        cr3[ci(i, k, j)] -= temp4 * ik(kz[k], cp[ci(i, k, j)]);
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cr2[ci(i, k, j)] -= temp4 * cs1[ci(i, k, j)];
  if (p.f_type == 1) {
    for (int j = js; j <= je; ++j) cr1[ci(0, 0, j)] -= temp4 * p.px0;
  } else if (p.f_type == 2) {
    for (int j = js; j <= je; ++j)
      cr1[ci(0, 0, j)] -= temp4 * (p.px0 + p.amp_omega0 * std::cos(p.omega0 * time));
  }
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        double s = -nu * std::pow(kx2[i] + kz2[k], p.beta);
        cf1[ci(i, k, j)] = s * cu1[ci(i, k, j)];
        cf3[ci(i, k, j)] = s * cu3[ci(i, k, j)];
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i)
        cf2[ci(i, k, j)] = (-nu * std::pow(kx2[i] + kz2[k], p.beta)) * cu2[ci(i, k, j)];
  // Coriolis (channel.f90:146-153)
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        cf1[ci(i, k, j)] += (p.ro_inv / p.delta) * cu3[ci(i, k, j)];
        cf3[ci(i, k, j)] -= (p.ro_inv / p.delta) * cu1[ci(i, k, j)];
      }
  for (int n = 0; n < Nth; ++n) {
    cplx* cth = CTH(th, n);
    cplx* cfth = CTH(fth, n);
    cplx* crth = CTH(rth, n);
    // buoyancy, skipping mode (0,0) (channel.f90:162-182)
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = (k == 0 ? 1 : 0); i < Nxp; ++i)
          cf2[ci(i, k, j)] += p.grav_y * (cth[ci(i, k, j)] * dyf[j - 1] + cth[ci(i, k, j - 1)] * dyf[j]) /
                              (2.0 * dy[j]);
#pragma omp parallel for schedule(static)
    for (int j = js; j <= je; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          cf1[ci(i, k, j)] += p.grav_x * cth[ci(i, k, j)];
          cf3[ci(i, k, j)] += p.grav_z * cth[ci(i, k, j)];
        }
    const int jst = jstart_th[n], jet = jend_th[n];
#pragma omp parallel for schedule(static)
    for (int j = jst; j <= jet; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          crth[ci(i, k, j)] = cth[ci(i, k, j)];
          if (rk_step > 1) crth[ci(i, k, j)] += temp3 * cfth[ci(i, k, j)];
          cfth[ci(i, k, j)] = -(nu / p.pr[n]) * std::pow(kx2[i] + kz2[k], p.beta) * cth[ci(i, k, j)];
          cfth[ci(i, k, j)] = cfth[ci(i, k, j)] - cu1[ci(i, k, j)] * p.dthdx[n] - cu3[ci(i, k, j)] * p.dthdz[n];
        }
  }
  // user_rhs_chan_fourier (forcing.f90:37-113): the thermal-wind advection of
  // the Front flavour (:50-100); a no-op otherwise
  if (p.flavor_front && p.ro_inv != 0.0) {
    const double Ly = p.ly, dl = p.delta, ro = p.ro_inv;
    for (int n = 0; n < Nth; ++n) {
      const double dx = p.dthdx[n], dz = p.dthdz[n];
      cplx* cth = CTH(th, n);
      cplx* cfth = CTH(fth, n);
#pragma omp parallel for schedule(static)
      for (int j = js; j <= je; ++j)  // forcing.f90:55-72
        for (int k = 0; k <= twoNkz; ++k)
          for (int i = 0; i < Nxp; ++i) {
            const cplx u2m = 0.5 * (cu2[ci(i, k, j)] + cu2[ci(i, k, j + 1)]);
            cf1[ci(i, k, j)] = cf1[ci(i, k, j)] - (dx * (gyf[j] - 0.5 * Ly) * dl / ro) * ik(kz[k], cu1[ci(i, k, j)]) -
                               (-1.0 * dz * (gyf[j] - 0.5 * Ly) * dl / ro) * ik(kx[i], cu1[ci(i, k, j)]) -
                               (-1.0 * dz * dl / ro) * u2m;
            cf3[ci(i, k, j)] = cf3[ci(i, k, j)] - (dx * (gyf[j] - 0.5 * Ly) * dl / ro) * ik(kz[k], cu3[ci(i, k, j)]) -
                               (-1.0 * dz * (gyf[j] - 0.5 * Ly) * dl / ro) * ik(kx[i], cu3[ci(i, k, j)]) -
                               (dx * dl / ro) * u2m;
          }
#pragma omp parallel for schedule(static)
      for (int j = 2; j <= Nyp; ++j)  // :74-84
        for (int k = 0; k <= twoNkz; ++k)
          for (int i = 0; i < Nxp; ++i)
            cf2[ci(i, k, j)] = cf2[ci(i, k, j)] - (dx * (gy[j] - 0.5 * Ly) * dl / ro) * ik(kz[k], cu2[ci(i, k, j)]) -
                               (-1.0 * dz * (gy[j] - 0.5 * Ly) * dl / ro) * ik(kx[i], cu2[ci(i, k, j)]);
#pragma omp parallel for schedule(static)
      for (int j = jstart_th[n]; j <= jend_th[n]; ++j)  // :87-96
        for (int k = 0; k <= twoNkz; ++k)
          for (int i = 0; i < Nxp; ++i)
            cfth[ci(i, k, j)] = cfth[ci(i, k, j)] - (dl / ro) * dx * (gyf[j] - 0.5 * Ly) * ik(kz[k], cth[ci(i, k, j)]) -
                                (dl / ro) * (-1.0) * dz * (gyf[j] - 0.5 * Ly) * ik(kx[i], cth[ci(i, k, j)]);
    }
  }

  // ---- LES gate (channel.f90:259) or plain inverse FFTs (:279-287)
  if (p.use_les && (!p.create_new_flow || time_step > 100)) {
    les_chan(time);
    les_chan_th(time);
  } else {
    fft_xz_to_physical(u1.data());
    fft_xz_to_physical(u2.data());
    fft_xz_to_physical(u3.data());
    for (int n = 0; n < Nth; ++n) fft_xz_to_physical(TH(th, n));
  }

  // ---- nonlinear products (channel.f90:299-459)
  double* S = s1.data();
  double* U1 = u1.data();
  double* U2 = u2.data();
  double* U3 = u3.data();
  auto prod = [&](int ja, int jb, auto f) {
#pragma omp parallel for schedule(static)
    for (int j = ja; j <= jb; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) S[ri(i, k, j)] = f(i, k, j);
  };
  auto acc = [&](int ja, int jb, auto f) {
#pragma omp parallel for schedule(static)
    for (int j = ja; j <= jb; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) f(i, k, j);
  };
  prod(js, je, [&](int i, int k, int j) { return U3[ri(i, k, j)] * U1[ri(i, k, j)]; });
  fft_xz_to_fourier(S);
  acc(js, je, [&](int i, int k, int j) {
    cf1[ci(i, k, j)] -= ik(kz[k], cs1[ci(i, k, j)]);
    cf3[ci(i, k, j)] -= ik(kx[i], cs1[ci(i, k, j)]);
  });
  prod(js, je, [&](int i, int k, int j) { return U1[ri(i, k, j)] * U1[ri(i, k, j)]; });
  fft_xz_to_fourier(S);
  acc(js, je, [&](int i, int k, int j) { cf1[ci(i, k, j)] -= ik(kx[i], cs1[ci(i, k, j)]); });
  prod(js, je, [&](int i, int k, int j) { return U3[ri(i, k, j)] * U3[ri(i, k, j)]; });
  fft_xz_to_fourier(S);
  acc(js, je, [&](int i, int k, int j) { cf3[ci(i, k, j)] -= ik(kz[k], cs1[ci(i, k, j)]); });
  prod(2, Nyp, [&](int i, int k, int j) {
    return ((dyf[j] * U1[ri(i, k, j)] + dyf[j - 1] * U1[ri(i, k, j - 1)]) / (2.0 * dy[j])) *
           U2[ri(i, k, j)];
  });
  fft_xz_to_fourier(S);
  acc(2, Nyp, [&](int i, int k, int j) { cf2[ci(i, k, j)] -= ik(kx[i], cs1[ci(i, k, j)]); });
  prod(2, Nyp, [&](int i, int k, int j) {
    return ((dyf[j] * U3[ri(i, k, j)] + dyf[j - 1] * U3[ri(i, k, j - 1)]) / (2.0 * dy[j])) *
           U2[ri(i, k, j)];
  });
  fft_xz_to_fourier(S);
  acc(2, Nyp, [&](int i, int k, int j) { cf2[ci(i, k, j)] -= ik(kz[k], cs1[ci(i, k, j)]); });
  prod(js, je, [&](int i, int k, int j) {
    return (U1[ri(i, k, j + 1)] * U2[ri(i, k, j + 1)] + U1[ri(i, k, j)] * U2[ri(i, k, j + 1)] -
            U1[ri(i, k, j)] * U2[ri(i, k, j)] - U1[ri(i, k, j - 1)] * U2[ri(i, k, j)]) /
           (2.0 * dyf[j]);
  });
  fft_xz_to_fourier(S);
  acc(js, je, [&](int i, int k, int j) { cf1[ci(i, k, j)] -= cs1[ci(i, k, j)]; });
  prod(js, je, [&](int i, int k, int j) {
    return (U3[ri(i, k, j + 1)] * U2[ri(i, k, j + 1)] + U3[ri(i, k, j)] * U2[ri(i, k, j + 1)] -
            U3[ri(i, k, j)] * U2[ri(i, k, j)] - U3[ri(i, k, j - 1)] * U2[ri(i, k, j)]) /
           (2.0 * dyf[j]);
  });
  fft_xz_to_fourier(S);
  acc(js, je, [&](int i, int k, int j) { cf3[ci(i, k, j)] -= cs1[ci(i, k, j)]; });
  prod(2, Nyp, [&](int i, int k, int j) {
    double a = U2[ri(i, k, j)] + U2[ri(i, k, j + 1)];
    double b = U2[ri(i, k, j)] + U2[ri(i, k, j - 1)];
    return (0.25 * (a * a) - 0.25 * (b * b)) / dy[j];
  });
  fft_xz_to_fourier(S);
  acc(2, Nyp, [&](int i, int k, int j) { cf2[ci(i, k, j)] -= cs1[ci(i, k, j)]; });

  // ---- cr += temp5*cf, to physical space (channel.f90:468-487)
  acc(js, je, [&](int i, int k, int j) {
    cr1[ci(i, k, j)] += temp5 * cf1[ci(i, k, j)];
    cr3[ci(i, k, j)] += temp5 * cf3[ci(i, k, j)];
  });
  acc(2, Nyp, [&](int i, int k, int j) { cr2[ci(i, k, j)] += temp5 * cf2[ci(i, k, j)]; });
  fft_xz_to_physical(r1.data());
  fft_xz_to_physical(r2.data());
  fft_xz_to_physical(r3.data());
  double* R1 = r1.data();
  double* R2 = r2.data();
  double* R3 = r3.data();
  double* NT = nu_t.data();

  // ---- explicit half of the vertical CN terms (channel.f90:494-547)
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k < Nzp; ++k)
      for (int i = 0; i < Nx; ++i) {
        size_t o = ri(i, k, j), up = ri(i, k, j + 1), dn = ri(i, k, j - 1);
        R1[o] += temp1 * nvs * (((U1[up] - U1[o]) / dy[j + 1] - (U1[o] - U1[dn]) / dy[j]) / dyf[j]);
        R3[o] += temp1 * nvs * (((U3[up] - U3[o]) / dy[j + 1] - (U3[o] - U3[dn]) / dy[j]) / dyf[j]);
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k < Nzp; ++k)
      for (int i = 0; i < Nx; ++i) {
        size_t o = ri(i, k, j), up = ri(i, k, j + 1), dn = ri(i, k, j - 1);
        R2[o] += temp1 * nvs * (((U2[up] - U2[o]) / dyf[j] - (U2[o] - U2[dn]) / dyf[j - 1]) / dy[j]);
      }
  if (p.use_les) {
#pragma omp parallel for schedule(static)
    for (int j = js; j <= je; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          size_t o = ri(i, k, j), up = ri(i, k, j + 1), dn = ri(i, k, j - 1);
          R1[o] += temp2 * ((NT[up] * (U1[up] - U1[o]) / dy[j + 1] - NT[o] * (U1[o] - U1[dn]) / dy[j]) / dyf[j]);
          R3[o] += temp2 * ((NT[up] * (U3[up] - U3[o]) / dy[j + 1] - NT[o] * (U3[o] - U3[dn]) / dy[j]) / dyf[j]);
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          size_t o = ri(i, k, j), up = ri(i, k, j + 1), dn = ri(i, k, j - 1);
          R2[o] += temp2 * ((0.5 * (NT[o] + NT[up]) * (U2[up] - U2[o]) / dyf[j] -
                             0.5 * (NT[o] + NT[dn]) * (U2[o] - U2[dn]) / dyf[j - 1]) /
                            dy[j]);
        }
  }

  // ---- scalars: products, RHS, implicit solve (channel.f90:551-711)
  for (int n = 0; n < Nth; ++n) {
    double* THn = TH(th, n);
    double* RTHn = TH(rth, n);
    double* KT = TH(kappa_t, n);
    cplx* cfth = CTH(fth, n);
    cplx* crth = CTH(rth, n);
    const int jst = jstart_th[n], jet = jend_th[n];
    prod(jst, jet, [&](int i, int k, int j) { return THn[ri(i, k, j)] * U1[ri(i, k, j)]; });
    fft_xz_to_fourier(S);
    acc(jst, jet, [&](int i, int k, int j) { cfth[ci(i, k, j)] -= ik(kx[i], cs1[ci(i, k, j)]); });
    prod(jst, jet, [&](int i, int k, int j) { return THn[ri(i, k, j)] * U3[ri(i, k, j)]; });
    fft_xz_to_fourier(S);
    acc(jst, jet, [&](int i, int k, int j) { cfth[ci(i, k, j)] -= ik(kz[k], cs1[ci(i, k, j)]); });
    prod(jst, jet, [&](int i, int k, int j) {
      return (THn[ri(i, k, j + 1)] * U2[ri(i, k, j + 1)] + THn[ri(i, k, j)] * U2[ri(i, k, j + 1)] -
              THn[ri(i, k, j)] * U2[ri(i, k, j)] - THn[ri(i, k, j - 1)] * U2[ri(i, k, j)]) /
             (2.0 * dyf[j]);
    });
    fft_xz_to_fourier(S);
    acc(jst, jet, [&](int i, int k, int j) { cfth[ci(i, k, j)] -= cs1[ci(i, k, j)]; });
    acc(jst, jet, [&](int i, int k, int j) { crth[ci(i, k, j)] += temp5 * cfth[ci(i, k, j)]; });
    fft_xz_to_physical(RTHn);
    const double tp = temp1 / p.pr[n];
#pragma omp parallel for schedule(static)
    for (int j = jst; j <= jet; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          size_t o = ri(i, k, j), up = ri(i, k, j + 1), dn = ri(i, k, j - 1);
          RTHn[o] += tp * nvs * (((THn[up] - THn[o]) / dy[j + 1] - (THn[o] - THn[dn]) / dy[j]) / dyf[j]);
          if (p.use_les)
            RTHn[o] += temp2 * ((KT[up] * (THn[up] - THn[o]) / dy[j + 1] - KT[o] * (THn[o] - THn[dn]) / dy[j]) / dyf[j]);
        }
#pragma omp parallel
    {
      std::vector<double> ml((size_t)Nx * NY2), md((size_t)Nx * NY2), mu((size_t)Nx * NY2), vec((size_t)Nx * NY2);
#pragma omp for schedule(static)
      for (int k = 0; k < Nzp; ++k) {
        for (size_t o = 0; o < ml.size(); ++o) { ml[o] = 0.0; md[o] = 1.0; mu[o] = 0.0; vec[o] = 0.0; }
        for (int j = jst; j <= jet; ++j)
          for (int i = 0; i < Nx; ++i) {
            size_t o = i + (size_t)Nx * j;
            ml[o] = -(tp * nvs) / (dy[j] * dyf[j]);
            md[o] = 1.0 + (tp * nvs) / (dy[j + 1] * dyf[j]) + (tp * nvs) / (dy[j] * dyf[j]);
            mu[o] = -(tp * nvs) / (dy[j + 1] * dyf[j]);
            vec[o] = RTHn[ri(i, k, j)];
          }
        if (p.use_les)
          for (int j = jst; j <= jet; ++j)
            for (int i = 0; i < Nx; ++i) {
              size_t o = i + (size_t)Nx * j;
              ml[o] -= temp2 * KT[ri(i, k, j)] / (dy[j] * dyf[j]);
              md[o] += temp2 * KT[ri(i, k, j + 1)] / (dy[j + 1] * dyf[j]) + temp2 * KT[ri(i, k, j)] / (dy[j] * dyf[j]);
              mu[o] -= temp2 * KT[ri(i, k, j + 1)] / (dy[j + 1] * dyf[j]);
            }
        apply_bc_rows(4 + n, ml.data(), md.data(), mu.data(), vec.data(), time);
        thomas_real(ml.data(), md.data(), mu.data(), vec.data(), Nyp, Nx);
        for (int j = jst - 1; j <= jet + 1; ++j)
          for (int i = 0; i < Nx; ++i) THn[ri(i, k, j)] = vec[i + (size_t)Nx * j];
      }
    }
  }

  // ---- velocity implicit solves: u2, u1, u3 (channel.f90:713-877)
  auto solve_vel = [&](int kind, double* U, double* Rr) {
    const bool isv = (kind == 2);
    const int ja = isv ? 2 : js, jb = isv ? Nyp : je;
#pragma omp parallel
    {
      std::vector<double> ml((size_t)Nx * NY2), md((size_t)Nx * NY2), mu((size_t)Nx * NY2), vec((size_t)Nx * NY2);
#pragma omp for schedule(static)
      for (int k = 0; k < Nzp; ++k) {
        for (size_t o = 0; o < ml.size(); ++o) { ml[o] = 0.0; md[o] = 1.0; mu[o] = 0.0; vec[o] = 0.0; }
        for (int j = ja; j <= jb; ++j)
          for (int i = 0; i < Nx; ++i) {
            size_t o = i + (size_t)Nx * j;
            if (isv) {
              ml[o] = -temp1 * nvs / (dyf[j - 1] * dy[j]);
              md[o] = 1.0 + temp1 * nvs / (dyf[j] * dy[j]) + temp1 * nvs / (dyf[j - 1] * dy[j]);
              mu[o] = -temp1 * nvs / (dyf[j] * dy[j]);
            } else {
              ml[o] = -temp1 * nvs / (dy[j] * dyf[j]);
              md[o] = 1.0 - temp1 * nvs * (-1.0 / (dy[j + 1] * dyf[j]) - 1.0 / (dy[j] * dyf[j]));
              mu[o] = -temp1 * nvs / (dy[j + 1] * dyf[j]);
            }
            vec[o] = Rr[ri(i, k, j)];
          }
        if (p.use_les)
          for (int j = ja; j <= jb; ++j)
            for (int i = 0; i < Nx; ++i) {
              size_t o = i + (size_t)Nx * j;
              if (isv) {
                double nl = 0.5 * (NT[ri(i, k, j)] + NT[ri(i, k, j - 1)]);
                double nu_ = 0.5 * (NT[ri(i, k, j)] + NT[ri(i, k, j + 1)]);
                ml[o] -= temp2 * nl / (dyf[j - 1] * dy[j]);
                md[o] += temp2 * nu_ / (dyf[j] * dy[j]) + temp2 * nl / (dyf[j - 1] * dy[j]);
                mu[o] -= temp2 * nu_ / (dyf[j] * dy[j]);
              } else {
                ml[o] -= temp2 * NT[ri(i, k, j)] / (dy[j] * dyf[j]);
                md[o] += temp2 * NT[ri(i, k, j + 1)] / (dy[j + 1] * dyf[j]) + temp2 * NT[ri(i, k, j)] / (dy[j] * dyf[j]);
                mu[o] -= temp2 * NT[ri(i, k, j + 1)] / (dy[j + 1] * dyf[j]);
              }
            }
        apply_bc_rows(kind, ml.data(), md.data(), mu.data(), vec.data(), time);
        thomas_real(ml.data(), md.data(), mu.data(), vec.data(), Nyp, Nx);
        const int ca = isv ? 1 : js - 1, cb = isv ? Nyp + 1 : je + 1;
        for (int j = ca; j <= cb; ++j)
          for (int i = 0; i < Nx; ++i) U[ri(i, k, j)] = vec[i + (size_t)Nx * j];
      }
    }
  };
  solve_vel(2, U2, R2);
  solve_vel(1, U1, R1);
  solve_vel(3, U3, R3);

  // variable_dt: courant on the velocity just solved (channel.f90:882-888)
  if (var_dt && rk_step == 3 && time_step % cour.update_dt == 0) courant(U1, U2, U3, time);

  // ---- back to Fourier space, projection, pressure (channel.f90:891-914)
  fft_xz_to_fourier(U1);
  fft_xz_to_fourier(U2);
  fft_xz_to_fourier(U3);
  for (int n = 0; n < Nth; ++n) fft_xz_to_fourier(TH(th, n));
  rem_div_chan();
#pragma omp parallel for schedule(static)
  for (int j = js; j <= je; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) cp[ci(i, k, j)] += cr1[ci(i, k, j)] / temp4;
}

}  // namespace orc

// ===================================================================== C-ABI
using orc::Solver;

struct OrcHandle {
  Solver* s;
  std::string err;
};

static int fail(OrcHandle* h, int code, const std::string& m) {
  if (h) h->err = m;
  return code;
}

static int check_params(const dgpu_params* p, std::string& why) {
  auto pow2 = [](int n) { return n >= 8 && (n & (n - 1)) == 0; };
  if (!pow2(p->nx)) { why = "nx must be a power of two >= 8"; return -1; }
  if (p->nz < 8) { why = "nz must be at least 8"; return -1; }
  if (p->ny < 5) { why = "ny must be >= 5"; return -1; }
  if (p->n_th < 0 || p->n_th > DGPU_MAX_TH) { why = "n_th out of range"; return -1; }
  if (p->use_les && p->les_model_type != 1 && p->les_model_type != 4 && p->les_model_type != 5) {
    why = "les_model_type must be 1, 4 or 5";
    return -1;
  }
  if (!(p->re > 0) || !(p->delta_t > 0)) { why = "re and delta_t must be positive"; return -1; }
  return 0;
}

extern "C" {

int orc_create(const dgpu_params* p, const double* gy, dgpu_handle** out) {
  if (!p || !gy || !out) return DGPU_EINVAL;
  std::string why;
  if (check_params(p, why)) return DGPU_EINVAL;
  OrcHandle* h = new (std::nothrow) OrcHandle();
  if (!h) return DGPU_ENOMEM;
  try {
    h->s = new Solver(*p, gy);
  } catch (...) {
    delete h;
    return DGPU_ENOMEM;
  }
  *out = reinterpret_cast<dgpu_handle*>(h);
  return 0;
}

int orc_destroy(dgpu_handle* hh) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  delete h->s;
  delete h;
  return 0;
}

const char* orc_last_error(const dgpu_handle* hh) {
  const OrcHandle* h = reinterpret_cast<const OrcHandle*>(hh);
  return h ? h->err.c_str() : "null handle";
}

static std::vector<double>* field_ptr(Solver* s, int f, int64_t* n) {
  *n = (int64_t)s->NB;
  switch (f) {
    case DGPU_CU1: return &s->u1;
    case DGPU_CU2: return &s->u2;
    case DGPU_CU3: return &s->u3;
    case DGPU_CP: return &s->pp;
    case DGPU_CTH: *n *= s->Nth; return &s->th;
    case DGPU_CF1: return &s->f1;
    case DGPU_CF2: return &s->f2;
    case DGPU_CF3: return &s->f3;
    case DGPU_CFTH: *n *= s->Nth; return &s->fth;
    case DGPU_NU_T: return &s->nu_t;
    case DGPU_KAPPA_T: *n *= s->Nth; return &s->kappa_t;
  }
  return nullptr;
}

int64_t orc_field_doubles(const dgpu_handle* hh, int field) {
  const OrcHandle* h = reinterpret_cast<const OrcHandle*>(hh);
  if (!h) return -1;
  int64_t n;
  if (!field_ptr(h->s, field, &n)) return -1;
  return n;
}

int orc_set_field(dgpu_handle* hh, int field, const double* host) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !host) return DGPU_EINVAL;
  int64_t n;
  std::vector<double>* v = field_ptr(h->s, field, &n);
  if (!v) return fail(h, DGPU_EINVAL, "bad field id");
  std::memcpy(v->data(), host, sizeof(double) * n);
  return 0;
}

// Copies the buffer out; entries outside the dealiased block (i>Nkx or
// k>2Nkz of a spectral field) are written as zero.
int orc_get_field(dgpu_handle* hh, int field, double* host) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !host) return DGPU_EINVAL;
  int64_t n;
  std::vector<double>* v = field_ptr(h->s, field, &n);
  if (!v) return fail(h, DGPU_EINVAL, "bad field id");
  std::memcpy(host, v->data(), sizeof(double) * n);
  Solver* s = h->s;
  if (field != DGPU_NU_T && field != DGPU_KAPPA_T) {
    int64_t nf = n / (int64_t)s->NB;
    for (int64_t f = 0; f < nf; ++f) {
      double* b = host + f * s->NB;
      for (int j = 0; j < s->NY2; ++j)
        for (int k = 0; k < s->NZ2; ++k)
          for (int i = 0; i < s->NXC; ++i)
            if (i > s->Nkx || k > s->twoNkz) {
              size_t o = 2 * s->ci(i, k, j);
              b[o] = b[o + 1] = 0.0;
            }
    }
  }
  return 0;
}

int orc_apply_bc_vel_post(dgpu_handle* hh, double time) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  h->s->apply_bc_vel_post(time);
  return 0;
}

int orc_rem_div(dgpu_handle* hh) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  h->s->rem_div_chan();
  return 0;
}

int orc_rk_substep(dgpu_handle* hh, int rk_step, double time, int time_step) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || rk_step < 1 || rk_step > 3) return DGPU_EINVAL;
  h->s->rk_chan_1(rk_step, time, time_step);
  return 0;
}

int orc_step(dgpu_handle* hh, int nsteps, double* time, int32_t* time_step) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !time || !time_step || nsteps < 0) return DGPU_EINVAL;
  for (int s = 0; s < nsteps; ++s) {
    *time_step += 1;
    for (int rk = 1; rk <= 3; ++rk) h->s->rk_chan_1(rk, *time, *time_step);
    *time += h->s->p.delta_t;
  }
  return 0;
}

int orc_synchronize(dgpu_handle*) { return 0; }

// ---- test hooks on raw reference-layout buffers
int orc_fft_xz_to_fourier(dgpu_handle* hh, double* buf) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !buf) return DGPU_EINVAL;
  h->s->fft_xz_to_fourier(buf);
  return 0;
}
int orc_fft_xz_to_physical(dgpu_handle* hh, double* buf) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !buf) return DGPU_EINVAL;
  h->s->fft_xz_to_physical(buf);
  return 0;
}
// Physical copies of cu1..3, cp, cth around the flow-file I/O
// (phdf5.f90:1076-1082, 1190-1195, 1431-1446): fft_xz_to_physical of a copy,
// and fft_xz_to_fourier of the given physical field into the state.
static bool physical_ok(int field) {
  return field == DGPU_CU1 || field == DGPU_CU2 || field == DGPU_CU3 || field == DGPU_CP || field == DGPU_CTH;
}
int orc_get_physical(dgpu_handle* hh, int field, double* host) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !host) return DGPU_EINVAL;
  int64_t n;
  std::vector<double>* v = field_ptr(h->s, field, &n);
  if (!v || !physical_ok(field) || n == 0) return fail(h, DGPU_EINVAL, "no physical copy of this field");
  std::memcpy(host, v->data(), sizeof(double) * n);
  const int64_t NB = (int64_t)h->s->NB;
  for (int64_t f = 0; f < n / NB; ++f) h->s->fft_xz_to_physical(host + f * NB);
  return 0;
}
int orc_set_physical(dgpu_handle* hh, int field, const double* host) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !host) return DGPU_EINVAL;
  int64_t n;
  std::vector<double>* v = field_ptr(h->s, field, &n);
  if (!v || !physical_ok(field) || n == 0) return fail(h, DGPU_EINVAL, "no physical copy of this field");
  std::memcpy(v->data(), host, sizeof(double) * n);
  const int64_t NB = (int64_t)h->s->NB;
  for (int64_t f = 0; f < n / NB; ++f) h->s->fft_xz_to_fourier(v->data() + f * NB);
  return 0;
}
int orc_poisson_p(dgpu_handle* hh, const double* ri) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || (h->s->Nth > 0 && !ri)) return DGPU_EINVAL;
  h->s->poisson_p_chan(ri);
  return 0;
}
int orc_set_courant(dgpu_handle* hh, const dgpu_courant_params* c) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  if (!c) {
    h->s->var_dt = false;
    return 0;
  }
  if (!(c->cfl > 0.0) || !(c->max_dt > 0.0) || c->update_dt < 1)
    return fail(h, DGPU_EINVAL, "courant needs cfl > 0, max_dt > 0 and update_dt >= 1");
  h->s->cour = *c;
  h->s->var_dt = true;
  return 0;
}
// courant as pre_first_step calls it: u1..u3 are the buffers of cu1..cu3
int orc_courant(dgpu_handle* hh, double time) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  if (!h->s->var_dt) return fail(h, DGPU_EINVAL, "orc_courant needs the settings of orc_set_courant");
  Solver* s = h->s;
  s->courant(s->u1.data(), s->u2.data(), s->u3.data(), time);
  return 0;
}
int orc_get_delta_t(dgpu_handle* hh, double* dt) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !dt) return DGPU_EINVAL;
  *dt = h->s->p.delta_t;
  return 0;
}
int orc_set_delta_t(dgpu_handle* hh, double dt) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h || !(dt > 0.0) || !std::isfinite(dt)) return DGPU_EINVAL;
  Solver* s = h->s;
  s->p.delta_t = dt;
  s->h_bar[0] = dt * (8.0 / 15.0);
  s->h_bar[1] = dt * (2.0 / 15.0);
  s->h_bar[2] = dt * (5.0 / 15.0);
  return 0;
}
// metrics: gyf, dy, dyf (Ny+2 each), kx (Nx/2), kz (2Nkz+1)
int orc_metrics(dgpu_handle* hh, double* gyf, double* dy, double* dyf, double* kx, double* kz) {
  OrcHandle* h = reinterpret_cast<OrcHandle*>(hh);
  if (!h) return DGPU_EINVAL;
  Solver* s = h->s;
  for (int j = 0; j < s->NY2; ++j) { gyf[j] = s->gyf[j]; dy[j] = s->dy[j]; dyf[j] = s->dyf[j]; }
  for (int i = 0; i < s->Nxp; ++i) kx[i] = s->kx[i];
  for (int k = 0; k <= s->twoNkz; ++k) kz[k] = s->kz[k];
  return 0;
}
// Raw Thomas sweeps on (inx, Nyp+2) row-major-by-j arrays, for brute-force tests.
int orc_thomas_real(double* a, double* b, double* c, double* g, int nyp, int inx) {
  orc::thomas_real(a, b, c, g, nyp, inx);
  return 0;
}
int orc_thomas_complex(double* a, double* b, double* c, double* g, int nyp, int inx) {
  orc::thomas_complex(a, b, c, reinterpret_cast<orc::cplx*>(g), nyp, inx);
  return 0;
}

}  // extern "C"
