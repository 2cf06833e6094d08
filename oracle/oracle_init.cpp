// oracle_init.cpp -- CPU restatement of poisson_p_chan and courant (test
// infrastructure only).
//
//   for/channel.f90  poisson_p_chan (1026-1234), apply_BC_poisson_mpi (2019-2055)
//   for/advance.f90  courant (47-158), called from pre_first_step (36-38) and
//                    from rk_chan_1 (channel.f90:882-888)
//
// poisson_p_chan works on scratch arrays instead of the RK carry cf1..cf3
// that the reference borrows (rk_step 1, the next substep, never reads them).
// Parity status: poisson_p_chan pinned by closed-form single-mode products and
// a dense solve of the same rows (tests/test_oracle_init.py); courant by the
// closed forms of its scalar rules and a brute-force minimum.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "oracle_internal.h"

namespace orc {

void Solver::poisson_p_chan(const double* ri_n) {
  cplx* cu1 = C(u1);
  cplx* cu2 = C(u2);
  cplx* cu3 = C(u3);
  cplx* cth = reinterpret_cast<cplx*>(th.data());
  cplx* cpp = C(pp);
  std::vector<double> a(NB, 0.0), b(NB, 0.0), c(NB, 0.0), s(NB, 0.0);
  cplx* ca = reinterpret_cast<cplx*>(a.data());
  cplx* cb = reinterpret_cast<cplx*>(b.data());
  cplx* cc = reinterpret_cast<cplx*>(c.data());
  cplx* cs = reinterpret_cast<cplx*>(s.data());
  auto modes = [&](auto fn) {
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) fn(i, k, j, ci(i, k, j));
  };
  auto points = [&](auto fn) {
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) fn(ri(i, k, j));
  };
  // diagonal terms (channel.f90:1041-1072)
  modes([&](int i, int k, int j, size_t o) {
    ca[o] = ik(kx[i], cu1[o]);
    cb[o] = (cu2[ci(i, k, j + 1)] - cu2[o]) / dyf[j];
    cc[o] = ik(kz[k], cu3[o]);
  });
  fft_xz_to_physical(a.data());
  fft_xz_to_physical(b.data());
  fft_xz_to_physical(c.data());
  points([&](size_t o) {
    a[o] = a[o] * a[o];
    b[o] = b[o] * b[o];
    c[o] = c[o] * c[o];
  });
  fft_xz_to_fourier(a.data());
  fft_xz_to_fourier(b.data());
  fft_xz_to_fourier(c.data());
  modes([&](int, int, int, size_t o) { cs[o] = ca[o] + cb[o] + cc[o]; });
  // the three off-diagonal products (channel.f90:1075-1172)
  for (int term = 0; term < 3; ++term) {
    modes([&](int i, int k, int j, size_t o) {
      if (term == 0) {
        ca[o] = (cu1[ci(i, k, j + 1)] - cu1[ci(i, k, j - 1)]) / (2.0 * dyf[j]);
        cb[o] = ik(kx[i] * 0.5, cu2[o] + cu2[ci(i, k, j + 1)]);
      } else if (term == 1) {
        ca[o] = (cu3[ci(i, k, j + 1)] - cu3[ci(i, k, j - 1)]) / (2.0 * dyf[j]);
        cb[o] = ik(kz[k] * 0.5, cu2[o] + cu2[ci(i, k, j + 1)]);
      } else {
        ca[o] = ik(kz[k], cu1[o]);
        cb[o] = ik(kx[i], cu3[o]);
      }
    });
    fft_xz_to_physical(a.data());
    fft_xz_to_physical(b.data());
    points([&](size_t o) { a[o] = 2.0 * a[o] * b[o]; });
    fft_xz_to_fourier(a.data());
    modes([&](int, int, int, size_t o) { cs[o] = cs[o] + ca[o]; });
  }
  // buoyancy, all modes but (0,0) (channel.f90:1178-1192)
  for (int n = 0; n < Nth; ++n) {
    const cplx* t = cth + (size_t)n * (NB / 2);
    modes([&](int i, int k, int j, size_t o) {
      if (i != 0 || k != 0)
        cs[o] = cs[o] + ri_n[n] * (t[ci(i, k, j + 1)] - t[ci(i, k, j - 1)]) / (gyf[j + 1] - gyf[j - 1]);
    });
  }
  // tridiagonal solve per kz (channel.f90:1196-1230)
  const int inx = Nxp;
  std::vector<double> ml((size_t)inx * NY2), md((size_t)inx * NY2), mu((size_t)inx * NY2);
  std::vector<cplx> vc((size_t)inx * NY2);
  for (int k = 0; k <= twoNkz; ++k) {
    for (int j = 0; j < NY2; ++j)
      for (int i = 0; i < inx; ++i) {
        const size_t o = i + (size_t)inx * j;
        ml[o] = 0.0;
        md[o] = 1.0;
        mu[o] = 0.0;
        vc[o] = cplx(0.0, 0.0);
      }
    for (int j = 2; j <= Nyp; ++j)
      for (int i = 0; i < inx; ++i) {
        const size_t o = i + (size_t)inx * j;
        ml[o] = 1.0 / (dy[j] * dyf[j]);
        md[o] = -kx2[i] - kz2[k] - 1.0 / (dy[j + 1] * dyf[j]) - 1.0 / (dy[j] * dyf[j]);
        mu[o] = 1.0 / (dy[j + 1] * dyf[j]);
        vc[o] = -1.0 * cs[ci(i, k, j)];
      }
    // apply_BC_poisson_mpi: row 1 Neumann (Dirichlet for the (0,0) mode);
    // row Nyp p(Nyp) = p(Nyp-1), its upper entry left as set above
    for (int i = 0; i < inx; ++i) {
      const size_t o = i + (size_t)inx;
      md[o] = 1.0;
      mu[o] = (i == 0 && k == 0) ? 0.0 : -1.0;
      vc[o] = cplx(0.0, 0.0);
      const size_t t = i + (size_t)inx * Nyp;
      md[t] = -1.0;
      ml[t] = 1.0;
      vc[t] = cplx(0.0, 0.0);
    }
    thomas_complex(ml.data(), md.data(), mu.data(), vc.data(), Nyp, inx);
    for (int j = 1; j <= Nyp; ++j)
      for (int i = 0; i < inx; ++i) cpp[ci(i, k, j)] = vc[i + (size_t)inx * j];
  }
}

// The smallest per-point limit of courant (advance.f90:110-133) over the
// real arrays a1..a3 in the physical layout (0:Nx+1, 0:Nz+1, 0:Ny+1).
double Solver::courant_points(const double* a1, const double* a2, const double* a3) {
  const double dx = p.lx / Nx, dz = p.lz / Nz;  // domain.f90:218, 260
  const double cfl = cour.cfl;
  double dt = HUGE_VAL;
  if (cour.front) {
    const double r = p.ro_inv / p.delta;
    for (int n = 0; n < Nth; ++n)
      for (int j = jstart; j <= jend; ++j)
        for (int k = 0; k < Nzp; ++k)
          for (int i = 0; i < Nx; ++i) {
            const size_t o = ri(i, k, j);
            const double dt_x = cfl * dx / std::fabs(a1[o] - p.dthdz[n] * gyf[j] / r);
            const double dt_y = cfl * dy[j] / std::fabs(a2[o]);
            const double dt_z = cfl * dz / std::fabs(a3[o] + (1.0 / r) * p.dthdx[n] * (gyf[j] - 0.5 * cour.ly));
            dt = std::min(std::min(dt, dt_x), std::min(dt_y, dt_z));
          }
  } else {
    for (int j = 1; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const size_t o = ri(i, k, j);
          const double dt_x = cfl * dx / std::fabs(a1[o]);
          const double dt_y = cfl * dy[j] / std::fabs(a2[o]);
          const double dt_z = cfl * dz / std::fabs(a3[o]);
          dt = std::min(std::min(dt, dt_x), std::min(dt_y, dt_z));
        }
  }
  return dt;
}

// courant (advance.f90:47-158): sets delta_t and h_bar.
void Solver::courant(const double* a1, const double* a2, const double* a3, double time) {
  double dt = cour.max_dt;
  if (p.ro_inv != 0.0) dt = std::min(dt, 2.0 * pi / std::fabs(p.ro_inv / p.delta) / 20.0);
  for (int n = 0; n < Nth; ++n) {
    const double nmax = std::sqrt(std::fabs(p.th_bc_ymin_c1[n]));
    dt = std::min(dt, (double)0.1f * 2.0 * pi / nmax);  // 0.1 is a default-real literal
  }
  dt = std::min(dt, courant_points(a1, a2, a3));
  double delta_t = dt <= 0.0 ? 0.00001 : (dt >= 1000.0 ? 1000.0 : dt);
  const double next = std::min(std::min(cour.time_limit, cour.save_stats_time - time),
                               std::min(cour.save_flow_time - time, cour.time_limit - time));
  const double r_max = (double)(1.0f - 0.025f);
  if (next < dt)
    delta_t = next + 1.0e-14;
  else if (next < (1.0 + r_max) * dt)
    delta_t = 1.0 / (1.0 + r_max) * next;
  else
    delta_t = dt;
  p.delta_t = delta_t;
  h_bar[0] = delta_t * (8.0 / 15.0);
  h_bar[1] = delta_t * (2.0 / 15.0);
  h_bar[2] = delta_t * (5.0 / 15.0);
}

}  // namespace orc
