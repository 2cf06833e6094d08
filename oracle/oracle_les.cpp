// oracle_les.cpp -- les_chan / les_chan_th (for/les.f90) on the CPU.
// TEST INFRASTRUCTURE ONLY (see oracle.cpp).
//
// Restates, at one y-rank (NprocY = NprocZ = 1), the subgrid models that
// for/les.f90 supports: constant Smagorinsky (type 1, :94-270), AMD-Rozema
// (type 4, :279-565) and AMD-Verstappen (type 5, :570-850), the stress
// divergence and explicit cross terms (:857-919), the scalar closures of
// les_chan_th (:1137-1415), and the helpers compute_strain_chan (:1463-1510),
// compute_strain_chan_invariant (:1514-1577), compute_rotation_chan
// (:1584-1618), compute_rotation_chan_invariant (:1621-1675),
// compute_all_gradients_chan (:1681-1767), apply_BC_les (:2001-2058),
// apply_BC_th_les (:2065-2100), the NprocY = 1 branches of ghost_les_mpi,
// ghost_les_th_mpi and ghost_CS2_mpi (:2250-2265, :2417-2432, :2528-2534),
// and apply_BC_th_{lower,upper}_post (for/boundary.f90:728-821).
// The statistics blocks (flag_save_LES) are not part of the substep.
#include <cmath>

#include "oracle_internal.h"

namespace orc {

namespace {
enum LesBuf {
  SIJ1, SIJ2, SIJ3, SIJ4, SIJ5, SIJ6, TEMP, TEMP_TH, S1_TH,
  OIJ4, OIJ5, OIJ6,
  SI1, SI2, SI3, SI4, SI5, SI6,
  DU1DX, DU2DY, DU3DZ, DU2DX, DU3DX, DU1DZ, DU2DZ, DU1DY, DU3DY,
  NLES_FIXED
};
}  // namespace

double* Solver::lb(int id) {
  if (!les_ready) {
    les_buf.assign(NB * (NLES_FIXED + 3 * (Nth > 0 ? Nth : 1)), 0.0);
    les_ready = true;
  }
  return les_buf.data() + NB * (size_t)id;
}

// dthetadx/dy/dz of scalar n: ids after the fixed set
static inline int dth_id(int which, int n) { return NLES_FIXED + 3 * n + which; }

void Solver::apply_bc_th_post() {
  for (int n = 0; n < Nth; ++n) {
    cplx* cth = CTH(th, n);
    if (p.th_bc_ymin[n] == 0) {
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) cth[ci(i, k, 0)] = cth[ci(i, k, 1)] = 0.0;
      cth[ci(0, 0, 1)] = p.th_bc_ymin_c1[n];
      cth[ci(0, 0, 0)] = p.th_bc_ymin_c1[n];
    } else {
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          cth[ci(i, k, 1)] = cth[ci(i, k, 2)];
          cth[ci(i, k, 0)] = cth[ci(i, k, 1)];
        }
      cth[ci(0, 0, 1)] = cth[ci(0, 0, 2)] - dy[2] * p.th_bc_ymin_c1[n];
      cth[ci(0, 0, 0)] = cth[ci(0, 0, 1)] - dy[1] * p.th_bc_ymin_c1[n];
    }
    if (p.th_bc_ymax[n] == 0) {
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) cth[ci(i, k, Nyp)] = cth[ci(i, k, Nyp + 1)] = 0.0;
      cth[ci(0, 0, Nyp)] = p.th_bc_ymax_c1[n];
      cth[ci(0, 0, Nyp + 1)] = p.th_bc_ymax_c1[n];
    } else {
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          cth[ci(i, k, Nyp)] = cth[ci(i, k, Nyp - 1)];
          cth[ci(i, k, Nyp + 1)] = cth[ci(i, k, Nyp)];
        }
      cth[ci(0, 0, Nyp)] = cth[ci(0, 0, Nyp - 1)] + dy[Nyp] * p.th_bc_ymax_c1[n];
      cth[ci(0, 0, Nyp + 1)] = cth[ci(0, 0, Nyp)] + dy[Nyp] * p.th_bc_ymax_c1[n];
    }
  }
}

// apply_BC_les + apply_BC_th_les: Neumann walls copy the interior into the ghosts
void Solver::apply_bc_les() {
  cplx* cu1 = C(u1);
  cplx* cu3 = C(u3);
  auto copy_up = [&](cplx* c) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        c[ci(i, k, Nyp)] = c[ci(i, k, Nyp - 1)];
        c[ci(i, k, Nyp + 1)] = c[ci(i, k, Nyp)];
      }
  };
  auto copy_dn = [&](cplx* c) {
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        c[ci(i, k, 1)] = c[ci(i, k, 2)];
        c[ci(i, k, 0)] = c[ci(i, k, 1)];
      }
  };
  if (p.u_bc_ymax == 1) copy_up(cu1);
  if (p.w_bc_ymax == 1) copy_up(cu3);
  if (p.u_bc_ymin == 1) copy_dn(cu1);
  if (p.w_bc_ymin == 1) copy_dn(cu3);
  for (int n = 0; n < Nth; ++n) {
    if (p.th_bc_ymax[n] == 1) copy_up(CTH(th, n));
    if (p.th_bc_ymin[n] == 1) copy_dn(CTH(th, n));
  }
}

// compute_strain_chan (invariant = false) / compute_strain_chan_invariant
void Solver::compute_strain(bool invariant, double beta_sgs) {
  cplx *cu1 = C(u1), *cu2 = C(u2), *cu3 = C(u3);
  const int base = invariant ? SI1 : SIJ1;
  cplx* c[6];
  for (int m = 0; m < 6; ++m) c[m] = reinterpret_cast<cplx*>(lb(base + m));
  const double dxs = (p.lx / Nx) * beta_sgs, dzs = (p.lz / Nz) * beta_sgs;
  for (int j = 0; j <= Nyp + 1; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        c[0][o] = ik(kx[i], cu1[o]);
        c[2][o] = ik(kz[k], cu3[o]);
        if (invariant)
          c[4][o] = 0.5 * (ik(kz[k], cu1[o]) * (dzs / dxs) + ik(kx[i], cu3[o]) * (dxs / dzs));
        else
          c[4][o] = 0.5 * (ik(kz[k], cu1[o]) + ik(kx[i], cu3[o]));
        if (j != Nyp + 1) c[1][o] = (cu2[ci(i, k, j + 1)] - cu2[o]) / dyf[j];
      }
  for (int j = 1; j <= Nyp + 1; ++j) {
    const double dys = dy[j] * 2.0;
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j), m = ci(i, k, j - 1);
        if (invariant) {
          c[3][o] = 0.5 * ((cu1[o] - cu1[m]) / dy[j] * (dys / dxs) + ik(kx[i], cu2[o]) * (dxs / dys));
          c[5][o] = 0.5 * (ik(kz[k], cu2[o]) * (dzs / dys) + (cu3[o] - cu3[m]) / dy[j] * (dys / dzs));
        } else {
          c[3][o] = 0.5 * ((cu1[o] - cu1[m]) / dy[j] + ik(kx[i], cu2[o]));
          c[5][o] = 0.5 * (ik(kz[k], cu2[o]) + (cu3[o] - cu3[m]) / dy[j]);
        }
      }
  }
  // ghost_CS2_mpi at NprocY = 1: sentinel at the top ghost
  for (int k = 0; k <= twoNkz; ++k)
    for (int i = 0; i < Nxp; ++i) c[1][ci(i, k, Nyp + 1)] = cplx(-10000.0, 0.0);
  for (int m = 0; m < 6; ++m) fft_xz_to_physical(lb(base + m));
}

// compute_rotation_chan (invariant = false) / compute_rotation_chan_invariant
void Solver::compute_rotation(bool invariant, double beta_sgs) {
  cplx *cu1 = C(u1), *cu2 = C(u2), *cu3 = C(u3);
  cplx* o4 = reinterpret_cast<cplx*>(lb(OIJ4));
  cplx* o5 = reinterpret_cast<cplx*>(lb(OIJ5));
  cplx* o6 = reinterpret_cast<cplx*>(lb(OIJ6));
  const double dxs = (p.lx / Nx) * beta_sgs, dzs = (p.lz / Nz) * beta_sgs;
  for (int j = 0; j <= Nyp + 1; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        if (invariant)
          o5[o] = 0.5 * (ik(kz[k], cu1[o]) * (dzs / dxs) - ik(kx[i], cu3[o]) * (dxs / dzs));
        else
          o5[o] = 0.5 * (ik(kz[k], cu1[o]) - ik(kx[i], cu3[o]));
      }
  for (int j = 1; j <= Nyp + 1; ++j) {
    const double dys = dy[j] * 2.0;
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j), m = ci(i, k, j - 1);
        if (invariant) {
          o4[o] = 0.5 * ((cu1[o] - cu1[m]) / dy[j] * (dys / dxs) - ik(kx[i], cu2[o]) * (dxs / dys));
          o6[o] = 0.5 * (ik(kz[k], cu2[o]) * (dzs / dys) - (cu3[o] - cu3[m]) / dy[j] * (dys / dzs));
        } else {
          o4[o] = 0.5 * ((cu1[o] - cu1[m]) / dy[j] - ik(kx[i], cu2[o]));
          o6[o] = 0.5 * (ik(kz[k], cu2[o]) - (cu3[o] - cu3[m]) / dy[j]);
        }
      }
  }
  fft_xz_to_physical(lb(OIJ4));
  fft_xz_to_physical(lb(OIJ5));
  fft_xz_to_physical(lb(OIJ6));
}

// compute_all_gradients_chan
void Solver::compute_all_gradients() {
  cplx *cu1 = C(u1), *cu2 = C(u2), *cu3 = C(u3);
  auto cb = [&](int id) { return reinterpret_cast<cplx*>(lb(id)); };
  for (int j = 0; j <= Nyp + 1; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        cb(DU1DX)[o] = ik(kx[i], cu1[o]);
        cb(DU1DZ)[o] = ik(kz[k], cu1[o]);
        cb(DU2DX)[o] = ik(kx[i], cu2[o]);
        cb(DU2DZ)[o] = ik(kz[k], cu2[o]);
        cb(DU3DX)[o] = ik(kx[i], cu3[o]);
        cb(DU3DZ)[o] = ik(kz[k], cu3[o]);
      }
  for (int j = 1; j <= Nyp + 1; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j), m = ci(i, k, j - 1);
        cb(DU1DY)[o] = (cu1[o] - cu1[m]) / dy[j];
        cb(DU2DY)[o] = (cu2[o] - cu2[m]) / dyf[j - 1];  // les.f90:1713, as written there
        cb(DU3DY)[o] = (cu3[o] - cu3[m]) / dy[j];
      }
  for (int n = 0; n < Nth; ++n) {
    cplx* cth = CTH(th, n);
    for (int j = 1; j <= Nyp + 1; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          const size_t o = ci(i, k, j);
          cb(dth_id(0, n))[o] = ik(kx[i], cth[o]);
          cb(dth_id(2, n))[o] = ik(kz[k], cth[o]);
          cb(dth_id(1, n))[o] = (cth[o] - cth[ci(i, k, j - 1)]) / dy[j];
        }
  }
  for (int id : {DU1DX, DU2DY, DU3DZ, DU2DX, DU3DX, DU1DZ, DU2DZ, DU1DY, DU3DY}) fft_xz_to_physical(lb(id));
  for (int w = 0; w < 3; ++w)
    for (int n = 0; n < Nth; ++n) fft_xz_to_physical(lb(dth_id(w, n)));
}

int Solver::les_chan(double time) {
  const int type = p.les_model_type;
  if (type != 1 && type != 4 && type != 5) {
    err = "LES_MODEL_TYPE must be 1, 4 or 5 (les.f90:272-277)";
    return DGPU_EUNSUP;
  }
  const double c_smag = 0.17, c_amd = 0.2887, beta_sgs = 3.0;
  const double dx1 = p.lx / Nx, dz1 = p.lz / Nz;  // domain.f90:218, 260
  // j1, j2 at one y-rank: the rankY == NprocY-1 branch (les.f90:63-66)
  const int j1 = jstart, j2 = Nyp - 1;

  apply_bc_vel_post(time);
  apply_bc_th_post();
  apply_bc_les();

  double *S1 = lb(SIJ1), *S2 = lb(SIJ2), *S3 = lb(SIJ3), *S4 = lb(SIJ4), *S5 = lb(SIJ5),
         *S6 = lb(SIJ6), *TP = lb(TEMP);
  double *U1 = u1.data(), *U3 = u3.data(), *SS = s1.data();
  std::vector<double> delta_y(NY2, 0.0), delta_yf(NY2, 0.0);
  auto P = [&](double* a, int i, int k, int j) -> double& { return a[ri(i, k, j)]; };
  auto gy_interp = [&](double* a, int i, int k, int j) {  // GYF -> GY, les.f90:150-165
    return (a[ri(i, k, j)] * dyf[j - 1] + a[ri(i, k, j - 1)] * dyf[j]) / (2.0 * dy[j]);
  };
  auto avg_up = [&](double* a, int i, int k, int j) {  // GY -> GYF
    return 0.5 * (a[ri(i, k, j + 1)] + a[ri(i, k, j)]);
  };

  if (type == 1) {
    compute_strain(false, beta_sgs);
    fft_xz_to_physical(u1.data());
    fft_xz_to_physical(u2.data());
    fft_xz_to_physical(u3.data());
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double s4 = avg_up(S4, i, k, j), s6 = avg_up(S6, i, k, j);
          P(SS, i, k, j) = std::sqrt(2.0 * std::pow(P(S1, i, k, j), 2) + 4.0 * std::pow(s4, 2) +
                                     4.0 * std::pow(P(S5, i, k, j), 2) + 2.0 * std::pow(P(S2, i, k, j), 2) +
                                     4.0 * std::pow(s6, 2) + 2.0 * std::pow(P(S3, i, k, j), 2));
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i)
          P(TP, i, k, j) = std::sqrt(2.0 * std::pow(gy_interp(S1, i, k, j), 2) + 4.0 * std::pow(P(S4, i, k, j), 2) +
                                     4.0 * std::pow(gy_interp(S5, i, k, j), 2) +
                                     2.0 * std::pow(gy_interp(S2, i, k, j), 2) + 4.0 * std::pow(P(S6, i, k, j), 2) +
                                     2.0 * std::pow(gy_interp(S3, i, k, j), 2));
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S1, i, k, j) *= P(SS, i, k, j);
          P(S5, i, k, j) *= P(SS, i, k, j);
          P(S2, i, k, j) = 0.0;
          P(S3, i, k, j) *= P(SS, i, k, j);
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S4, i, k, j) = P(TP, i, k, j) * (P(S4, i, k, j) - 0.5 * (P(U1, i, k, j) - P(U1, i, k, j - 1)) / dy[j]);
          P(S6, i, k, j) = P(TP, i, k, j) * (P(S6, i, k, j) - 0.5 * (P(U3, i, k, j) - P(U3, i, k, j - 1)) / dy[j]);
        }
    for (int j = 1; j <= Nyp; ++j)
      delta_yf[j] = -2.0 * std::pow(c_smag, 2) * std::pow(dx1 * beta_sgs * dyf[j] * 1.0 * dz1 * beta_sgs, 2.0 / 3.0);
    for (int j = 1; j <= Nyp + 1; ++j)
      delta_y[j] = -2.0 * std::pow(c_smag, 2) * std::pow(dx1 * beta_sgs * dy[j] * 2.0 * dz1 * beta_sgs, 2.0 / 3.0);
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) P(nu_t.data(), i, k, j) = -0.5 * delta_y[j] * P(TP, i, k, j);
  } else if (type == 4) {
    compute_strain(false, beta_sgs);
    compute_rotation(false, beta_sgs);
    fft_xz_to_physical(u1.data());
    fft_xz_to_physical(u2.data());
    fft_xz_to_physical(u3.data());
    double *O4 = lb(OIJ4), *O5 = lb(OIJ5), *O6 = lb(OIJ6);
    const double dxq = std::pow(dx1 * beta_sgs, 2), dzq = std::pow(dz1 * beta_sgs, 2);
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp; ++j) {
      const double dyq = std::pow(dy[j] * 2.0, 2);  // les.f90:306 (dy, not dyf)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double a1 = P(S1, i, k, j), a2 = P(S2, i, k, j), a3 = P(S3, i, k, j);
          const double a4 = avg_up(S4, i, k, j), a5 = P(S5, i, k, j), a6 = avg_up(S6, i, k, j);
          const double w4 = avg_up(O4, i, k, j), w5 = P(O5, i, k, j), w6 = avg_up(O6, i, k, j);
          double v = dxq * std::pow(a1, 3) + dyq * std::pow(a2, 3) + dzq * std::pow(a3, 3) +
                     (2 * dxq + dyq) * a1 * a4 * a4 + (2 * dxq + dzq) * a1 * a5 * a5 +
                     (2 * dyq + dxq) * a2 * a4 * a4 + (2 * dyq + dzq) * a2 * a6 * a6 +
                     (2 * dzq + dxq) * a3 * a5 * a5 + (2 * dzq + dyq) * a3 * a6 * a6 +
                     2 * (dxq + dyq + dzq) * a4 * a5 * a6;
          v = -v - dyq * a1 * w4 * w4 - dzq * a1 * w5 * w5 - dxq * a2 * w4 * w4 - dzq * a2 * w6 * w6 -
              dxq * a3 * w5 * w5 - dyq * a3 * w6 * w6 - 2 * dzq * a4 * w5 * w6 + 2 * dyq * a5 * w4 * w6 -
              2 * dxq * a6 * w4 * w5;
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 +
                     2 * a6 * a6 + 2 * w6 * w6);
          P(SS, i, k, j) = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp + 1; ++j) {
      const double dyq = std::pow(dy[j] * 2.0, 2);
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double a1 = gy_interp(S1, i, k, j), a2 = gy_interp(S2, i, k, j), a3 = gy_interp(S3, i, k, j);
          const double a4 = P(S4, i, k, j), a5 = gy_interp(S5, i, k, j), a6 = P(S6, i, k, j);
          const double w4 = P(O4, i, k, j), w5 = gy_interp(O5, i, k, j), w6 = P(O6, i, k, j);
          double v = dxq * std::pow(a1, 3) + dyq * std::pow(a2, 3) + dzq * std::pow(a3, 3) +
                     (2 * dxq + dyq) * a1 * a4 * a4 + (2 * dxq + dzq) * a1 * a5 * a5 +
                     (2 * dyq + dxq) * a2 * a4 * a4 + (2 * dyq + dzq) * a2 * a6 * a6 +
                     (2 * dzq + dxq) * a3 * a5 * a5 + (2 * dzq + dyq) * a3 * a6 * a6 +
                     2 * (dxq + dyq + dzq) * a4 * a6 * a5;
          v = -v - dyq * a1 * w4 * w4 - dzq * a1 * w5 * w5 - dxq * a2 * w4 * w4 - dzq * a2 * w6 * w6 -
              dxq * a3 * w5 * w5 - dyq * a3 * w6 * w6 - 2 * dzq * a4 * w5 * w6 + 2 * dyq * a5 * w4 * w6 -
              2 * dxq * a6 * w4 * w5;
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 +
                     2 * a6 * a6 + 2 * w6 * w6);
          P(TP, i, k, j) = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S1, i, k, j) *= P(SS, i, k, j);
          P(S5, i, k, j) *= P(SS, i, k, j);
          P(S2, i, k, j) = 0.0;
          P(S3, i, k, j) *= P(SS, i, k, j);
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S4, i, k, j) = P(TP, i, k, j) * (P(S4, i, k, j) - 0.5 * (P(U1, i, k, j) - P(U1, i, k, j - 1)) / dy[j]);
          P(S6, i, k, j) = P(TP, i, k, j) * (P(S6, i, k, j) - 0.5 * (P(U3, i, k, j) - P(U3, i, k, j - 1)) / dy[j]);
        }
    for (int j = 1; j <= Nyp + 1; ++j) delta_yf[j] = delta_y[j] = -2.0 * std::pow(c_amd, 2);
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) P(nu_t.data(), i, k, j) = -0.5 * delta_y[j] * P(TP, i, k, j);
  } else {  // type 5, AMD-Verstappen
    compute_strain(false, beta_sgs);
    compute_strain(true, beta_sgs);
    compute_rotation(true, beta_sgs);
    compute_all_gradients();
    fft_xz_to_physical(u1.data());
    fft_xz_to_physical(u2.data());
    fft_xz_to_physical(u3.data());
    double *I1 = lb(SI1), *I2 = lb(SI2), *I3 = lb(SI3), *I4 = lb(SI4), *I5 = lb(SI5), *I6 = lb(SI6);
    double *O4 = lb(OIJ4), *O5 = lb(OIJ5), *O6 = lb(OIJ6);
#pragma omp parallel for schedule(static)
    for (int j = jstart; j <= jend; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double a1 = P(I1, i, k, j), a2 = P(I2, i, k, j), a3 = P(I3, i, k, j);
          const double a4 = avg_up(I4, i, k, j), a5 = P(I5, i, k, j), a6 = avg_up(I6, i, k, j);
          const double w4 = avg_up(O4, i, k, j), w5 = P(O5, i, k, j), w6 = avg_up(O6, i, k, j);
          double v = std::pow(a1, 3) + std::pow(a2, 3) + std::pow(a3, 3) + 3 * a1 * a4 * a4 + 3 * a1 * a5 * a5 +
                     3 * a2 * a4 * a4 + 3 * a2 * a6 * a6 + 3 * a3 * a5 * a5 + 3 * a3 * a6 * a6 +
                     6 * a4 * a5 * a6;
          v = -v - a1 * w4 * w4 - a1 * w5 * w5 - a2 * w4 * w4 - a2 * w6 * w6 - a3 * w5 * w5 -
              a3 * w6 * w6 - 2 * a4 * w5 * w6 + 2 * a5 * w4 * w6 - 2 * a6 * w4 * w5;
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 +
                     2 * a6 * a6 + 2 * w6 * w6);
          P(SS, i, k, j) = v;
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double a1 = gy_interp(I1, i, k, j), a2 = gy_interp(I2, i, k, j), a3 = gy_interp(I3, i, k, j);
          const double a4 = P(I4, i, k, j), a5 = gy_interp(I5, i, k, j), a6 = P(I6, i, k, j);
          const double w4 = P(O4, i, k, j), w5 = gy_interp(O5, i, k, j), w6 = P(O6, i, k, j);
          double v = std::pow(a1, 3) + std::pow(a2, 3) + std::pow(a3, 3) + 3 * a1 * a4 * a4 + 3 * a1 * a5 * a5 +
                     3 * a2 * a4 * a4 + 3 * a2 * a6 * a6 + 3 * a3 * a5 * a5 + 3 * a3 * a6 * a6 +
                     6 * a4 * a6 * a5;
          v = -v - a1 * w4 * w4 - a1 * w5 * w5 - a2 * w4 * w4 - a2 * w6 * w6 - a3 * w5 * w5 -
              a3 * w6 * w6 - 2 * a4 * w5 * w6 + 2 * a5 * w4 * w6 - 2 * a6 * w4 * w5;
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 +
                     2 * a6 * a6 + 2 * w6 * w6);
          P(TP, i, k, j) = v;
        }
#pragma omp parallel for schedule(static)
    for (int j = jstart; j <= jend; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S1, i, k, j) *= P(SS, i, k, j);
          P(S5, i, k, j) *= P(SS, i, k, j);
          P(S2, i, k, j) = 0.0;
          P(S3, i, k, j) *= P(SS, i, k, j);
        }
    // les.f90:757-769 runs to Nyp+1, where TEMP was not computed this call;
    // the value there is never consumed (nu_t at Nyp+1 is zeroed by the
    // ghost routine, CSij4/6 at Nyp+1 is not read), so TEMP's Nyp+1 plane
    // is treated as zero here.
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nzp; ++k)
      for (int i = 0; i < Nx; ++i) P(TP, i, k, Nyp + 1) = 0.0;
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(S4, i, k, j) = P(TP, i, k, j) * (P(S4, i, k, j) - 0.5 * (P(U1, i, k, j) - P(U1, i, k, j - 1)) / dy[j]);
          P(S6, i, k, j) = P(TP, i, k, j) * (P(S6, i, k, j) - 0.5 * (P(U3, i, k, j) - P(U3, i, k, j - 1)) / dy[j]);
        }
    for (int j = 1; j <= Nyp + 1; ++j) {
      delta_yf[j] = -2.0 * std::pow(c_amd, 2) * 3.0 /
                    (1.0 / std::pow(dx1 * beta_sgs, 2) + 1.0 / std::pow(dyf[j] * 2.0, 2) +
                     1.0 / std::pow(dz1 * beta_sgs, 2));
      delta_y[j] = -2.0 * std::pow(c_amd, 2) * 3.0 /
                   (1.0 / std::pow(dx1 * beta_sgs, 2) + 1.0 / std::pow(dy[j] * 2.0, 2) +
                    1.0 / std::pow(dz1 * beta_sgs, 2));
    }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp + 1; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) P(nu_t.data(), i, k, j) = -0.5 * delta_y[j] * P(TP, i, k, j);
  }
  // ghost_les_mpi at NprocY = 1 (les.f90:2250-2265)
#pragma omp parallel for schedule(static)
  for (int k = 0; k < Nzp; ++k)
    for (int i = 0; i < Nx; ++i)
      for (int j : {1, 2, Nyp, Nyp + 1}) P(nu_t.data(), i, k, j) = 0.0;

  for (double* b : {S1, S3, S4, S5, S6}) fft_xz_to_fourier(b);
  cplx *c1 = reinterpret_cast<cplx*>(S1), *c3 = reinterpret_cast<cplx*>(S3), *c4 = reinterpret_cast<cplx*>(S4),
       *c5 = reinterpret_cast<cplx*>(S5), *c6 = reinterpret_cast<cplx*>(S6);
  for (int j = 1; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        c1[o] = delta_yf[j] * c1[o];
        c5[o] = delta_yf[j] * c5[o];
        c3[o] = delta_yf[j] * c3[o];
      }
  for (int j = 1; j <= Nyp + 1; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        c4[o] = delta_y[j] * c4[o];
        c6[o] = delta_y[j] * c6[o];
      }
  // stress divergence into the RK terms (les.f90:861-885)
  cplx *cf1 = C(f1), *cf2 = C(f2), *cf3 = C(f3);
#pragma omp parallel for schedule(static)
  for (int j = j1; j <= j2; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j), u = ci(i, k, j + 1);
        cf1[o] = cf1[o] - ik(kx[i], c1[o]) - (c4[u] - c4[o]) / dyf[j] - ik(kz[k], c5[o]);
        cf3[o] = cf3[o] - ik(kx[i], c5[o]) - (c6[u] - c6[o]) / dyf[j] - ik(kz[k], c3[o]);
      }
#pragma omp parallel for schedule(static)
  for (int j = 2; j <= Nyp; ++j)
    for (int k = 0; k <= twoNkz; ++k)
      for (int i = 0; i < Nxp; ++i) {
        const size_t o = ci(i, k, j);
        cf2[o] = cf2[o] - ik(kx[i], c4[o]) - ik(kz[k], c6[o]);
      }
  // explicit cross terms (les.f90:890-919)
  double* NT = nu_t.data();
  for (int comp = 0; comp < 2; ++comp) {
    double* U = comp == 0 ? U1 : U3;
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) P(TP, i, k, j) = P(NT, i, k, j) * (P(U, i, k, j) - P(U, i, k, j - 1)) / dy[j];
    fft_xz_to_fourier(TP);
    cplx* ct = reinterpret_cast<cplx*>(TP);
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          const size_t o = ci(i, k, j);
          cf2[o] = cf2[o] + (comp == 0 ? ik(kx[i], ct[o]) : ik(kz[k], ct[o]));
        }
  }
  return 0;
}

int Solver::les_chan_th(double) {
  const int type = p.les_model_type;
  const double c_amd = 0.2887, beta_sgs = 3.0;
  const double dx1 = p.lx / Nx, dz1 = p.lz / Nz;
  auto P = [&](double* a, int i, int k, int j) -> double& { return a[ri(i, k, j)]; };
  if (type == 1 || type == 4) {
    // constant SGS Prandtl number, explicit horizontal diffusion (les.f90:1137-1209)
    for (int n = 0; n < Nth; ++n) {
      double* KT = TH(kappa_t, n);
#pragma omp parallel for schedule(static)
      for (int j = 1; j <= Nyp + 1; ++j)
        for (int k = 0; k < Nzp; ++k)
          for (int i = 0; i < Nx; ++i) P(KT, i, k, j) = 1.0 * P(nu_t.data(), i, k, j);
    }
    for (int n = 0; n < Nth; ++n) {
      double* KT = TH(kappa_t, n);
      cplx* cth = CTH(th, n);
      cplx* cfth = CTH(fth, n);
      cplx* cs1 = C(s1);
      for (int dir = 0; dir < 2; ++dir) {
        for (int j = 1; j <= Nyp + 1; ++j)
          for (int k = 0; k <= twoNkz; ++k)
            for (int i = 0; i < Nxp; ++i)
              cs1[ci(i, k, j)] = dir == 0 ? ik(kx[i], cth[ci(i, k, j)]) : ik(kz[k], cth[ci(i, k, j)]);
        fft_xz_to_physical(s1.data());
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= Nyp + 1; ++j)
          for (int k = 0; k < Nzp; ++k)
            for (int i = 0; i < Nx; ++i) P(s1.data(), i, k, j) = P(KT, i, k, j) * P(s1.data(), i, k, j);
        fft_xz_to_fourier(s1.data());
        for (int j = jstart_th[n]; j <= jend_th[n]; ++j)
          for (int k = 0; k <= twoNkz; ++k)
            for (int i = 0; i < Nxp; ++i) {
              const size_t o = ci(i, k, j);
              cfth[o] = cfth[o] + (dir == 0 ? ik(kx[i], cs1[o]) : ik(kz[k], cs1[o]));
            }
      }
    }
    for (int n = 0; n < Nth; ++n) fft_xz_to_physical(TH(th, n));
    return 0;
  }
  // type 5: AMD-Verstappen for the scalar (les.f90:1221-1415)
  for (int n = 0; n < Nth; ++n) fft_xz_to_physical(TH(th, n));
  const double deltax = dx1 * beta_sgs, deltaz = dz1 * beta_sgs;
  double *D11 = lb(DU1DX), *D22 = lb(DU2DY), *D33 = lb(DU3DZ), *D21 = lb(DU2DX), *D31 = lb(DU3DX),
         *D13 = lb(DU1DZ), *D23 = lb(DU2DZ), *D12 = lb(DU1DY), *D32 = lb(DU3DY);
  double *TT = lb(TEMP_TH), *ST = lb(S1_TH);
  for (int n = 0; n < Nth; ++n) {
    double *TX = lb(dth_id(0, n)), *TY = lb(dth_id(1, n)), *TZ = lb(dth_id(2, n));
    auto gyi = [&](double* a, int i, int k, int j) {
      return (a[ri(i, k, j)] * dyf[j - 1] + a[ri(i, k, j - 1)] * dyf[j]) / (2.0 * dy[j]);
    };
#pragma omp parallel for schedule(static)
    for (int j = jstart_th[n]; j <= jend_th[n]; ++j) {
      const double deltay = dyf[j] * 2.0;
      const double dx2 = deltax * deltax, dy2 = deltay * deltay, dz2 = deltaz * deltaz;
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double tx = P(TX, i, k, j), tz = P(TZ, i, k, j);
          const double ty = 0.5 * (P(TY, i, k, j + 1) + P(TY, i, k, j));
          double v = -(dx2 * P(D11, i, k, j) * tx * tx + dy2 * P(D22, i, k, j) * ty * ty +
                       dz2 * P(D33, i, k, j) * tz * tz +
                       (dy2 * 0.5 * (P(D12, i, k, j + 1) + P(D12, i, k, j)) +
                        dx2 * 0.5 * (P(D21, i, k, j + 1) + P(D21, i, k, j))) * tx * ty +
                       (dz2 * P(D13, i, k, j) + dx2 * P(D31, i, k, j)) * tx * tz +
                       (dz2 * 0.5 * (P(D23, i, k, j + 1) + P(D23, i, k, j)) +
                        dy2 * 0.5 * (P(D32, i, k, j + 1) + P(D32, i, k, j))) * ty * tz);
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (std::pow(deltax * tx, 2) + std::pow(deltay * ty, 2) + std::pow(deltaz * tz, 2));
          P(ST, i, k, j) = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp + 1; ++j) {
      const double deltay = dy[j] * 2.0;
      const double dx2 = deltax * deltax, dy2 = deltay * deltay, dz2 = deltaz * deltaz;
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          const double tx = gyi(TX, i, k, j), tz = gyi(TZ, i, k, j), ty = P(TY, i, k, j);
          double v = -(dx2 * gyi(D11, i, k, j) * tx * tx + dy2 * gyi(D22, i, k, j) * ty * ty +
                       dz2 * gyi(D33, i, k, j) * tz * tz +
                       (dy2 * P(D12, i, k, j) + dx2 * P(D21, i, k, j)) * tx * ty +
                       (dz2 * gyi(D13, i, k, j) + dx2 * gyi(D31, i, k, j)) * tx * tz +
                       (dz2 * P(D23, i, k, j) + dy2 * P(D32, i, k, j)) * ty * tz);
          if (v <= 0.0)
            v = 0.0;
          else
            v = v / (std::pow(deltax * tx, 2) + std::pow(deltay * ty, 2) + std::pow(deltaz * tz, 2));
          P(TT, i, k, j) = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = jstart_th[n]; j <= jend_th[n]; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) {
          P(TX, i, k, j) = P(ST, i, k, j) * P(TX, i, k, j);
          P(TZ, i, k, j) = P(ST, i, k, j) * P(TZ, i, k, j);
        }
    std::vector<double> delta_y(NY2, 0.0), delta_yf(NY2, 0.0);
    for (int j = 1; j <= Nyp + 1; ++j) {
      delta_yf[j] = -2.0 * std::pow(c_amd, 2) * 3.0 /
                    (1.0 / std::pow(dx1 * beta_sgs, 2) + 1.0 / std::pow(dyf[j] * 2.0, 2) +
                     1.0 / std::pow(dz1 * beta_sgs, 2));
      delta_y[j] = -2.0 * std::pow(c_amd, 2) * 3.0 /
                   (1.0 / std::pow(dx1 * beta_sgs, 2) + 1.0 / std::pow(dy[j] * 2.0, 2) +
                    1.0 / std::pow(dz1 * beta_sgs, 2));
    }
    double* KT = TH(kappa_t, n);
#pragma omp parallel for schedule(static)
    for (int j = 2; j <= Nyp; ++j)
      for (int k = 0; k < Nzp; ++k)
        for (int i = 0; i < Nx; ++i) P(KT, i, k, j) = -0.5 * delta_y[j] * P(TT, i, k, j);
    // ghost_les_th_mpi at NprocY = 1 (les.f90:2417-2432)
    for (int k = 0; k < Nzp; ++k)
      for (int i = 0; i < Nx; ++i)
        for (int j : {1, 2, Nyp, Nyp + 1}) P(KT, i, k, j) = 0.0;
    fft_xz_to_fourier(TX);
    fft_xz_to_fourier(TZ);
    cplx *cx = reinterpret_cast<cplx*>(TX), *cz = reinterpret_cast<cplx*>(TZ);
    for (int j = 1; j <= Nyp + 1; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          const size_t o = ci(i, k, j);
          cx[o] = 0.5 * delta_yf[j] * cx[o];
          cz[o] = 0.5 * delta_yf[j] * cz[o];
        }
    cplx* cfth = CTH(fth, n);
    for (int j = jstart_th[n]; j <= jend_th[n]; ++j)
      for (int k = 0; k <= twoNkz; ++k)
        for (int i = 0; i < Nxp; ++i) {
          const size_t o = ci(i, k, j);
          cfth[o] = cfth[o] - ik(kx[i], cx[o]) - ik(kz[k], cz[o]);
        }
  }
  return 0;
}

}  // namespace orc
