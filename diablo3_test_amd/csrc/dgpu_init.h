// Kernels of the start-up and time-step-control procedures around the
// substep: the pressure Poisson solve poisson_p_chan (channel.f90:1026-1234)
// and the CFL limit of courant (advance.f90:47-158).
#pragma once
#include "dgpu_kernels.h"

namespace dgpu {

// ------------------------------------------------------------- Poisson ----
// Pointwise part of the Poisson right-hand side in physical space,
// (du_i/dx_j)(du_j/dx_i) = A^2 + B^2 + C^2 + 2 D E + 2 F G + 2 H I
// (channel.f90:1041-1172), accumulated one term at a time on planes
// [j0, j1]: mode 0 acc = a*a, 1 acc += a*a, 2 acc += 2*a*b.  The reference
// transforms every term back separately; the transform is linear, so the
// sum is formed here and transformed once.
__global__ void __launch_bounds__(256) k_poisson_accum(double* __restrict__ acc, const double* __restrict__ a,
                                                       const double* __restrict__ b, int mode, int j0, int j1,
                                                       int nx, int nz, long long pp) {
  const long long n = (long long)(j1 - j0 + 1) * nz * nx;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const long long jl = t / ((long long)nz * nx);
    const long long r = t - jl * nz * nx;
    const long long o = (j0 + jl) * pp + r;  // P[j][k][i], i fastest, pitch nx
    const double x = a[o];
    if (mode == 0)
      acc[o] = x * x;
    else if (mode == 1)
      acc[o] = acc[o] + x * x;
    else
      acc[o] = acc[o] + 2.0 * x * b[o];
  }
}

struct PoissonArgs {
  const double2* rhs;  // spectral transform of the physical right-hand side (planes 2..Ny)
  const double2* cth[DGPU_MAX_TH];
  double ri[DGPU_MAX_TH];
  int nth;
  const double* kx;
  const double* kz;
  const double* dy;
  const double* dyf;
  const double* gyf;
  double* sb;   // Thomas scratch: main diagonal (Ny+2 planes of modes)
  double2* sg;  // Thomas scratch: right-hand side
  double2* cp;
};

// One lane per (kx, kz) mode: the buoyancy term (channel.f90:1178-1192, the
// (0,0) mode skipped), the right-hand side -cs1 and the complex tridiagonal
// solve of channel.f90:1196-1230 with the rows of apply_BC_poisson_mpi
// (channel.f90:2019-2055): row 1 Neumann (Dirichlet for the (0,0) mode), row
// Ny the Neumann row p(Ny) = p(Ny-1).  cp is written on planes 1..Ny.
__global__ void __launch_bounds__(256) k_poisson_solve(PoissonArgs a, Geo g) {
  typedef cpx<double> C;
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.sp) return;
  const int q = (int)((unsigned)m / (unsigned)g.nkk), kk = (int)((unsigned)m % (unsigned)g.nkk);
  const double kxq = a.kx[q], kzk = a.kz[kk];
  const double lam = kxq * kxq + kzk * kzk;
  const bool mean = (q == 0 && kk == 0);
  const int ny = g.ny;
  const long long sp = g.sp;
  const C* rhs = reinterpret_cast<const C*>(a.rhs);
  C* sg = reinterpret_cast<C*>(a.sg);
  C* cp = reinterpret_cast<C*>(a.cp);
  double bprev = 1.0;
  double muprev = mean ? 0.0 : -1.0;
  C gprev = {0.0, 0.0};
  a.sb[sp + m] = bprev;
  sg[sp + m] = gprev;
  for (int j = 2; j <= ny; ++j) {
    const long long o = (long long)j * sp + m;
    double ml, md, mu;
    C v;
    if (j < ny) {
      ml = 1.0 / (a.dy[j] * a.dyf[j]);
      md = -lam - 1.0 / (a.dy[j + 1] * a.dyf[j]) - 1.0 / (a.dy[j] * a.dyf[j]);
      mu = 1.0 / (a.dy[j + 1] * a.dyf[j]);
      C cs = rhs[o];
      if (!mean) {
        const double den = a.gyf[j + 1] - a.gyf[j - 1];
        for (int n = 0; n < a.nth; ++n) {
          const C* th = reinterpret_cast<const C*>(a.cth[n]);
          const C d = th[o + sp] - th[o - sp];
          cs = cs + C{a.ri[n] * d.x / den, a.ri[n] * d.y / den};
        }
      }
      v = C{-cs.x, -cs.y};
    } else {
      ml = 1.0;
      md = -1.0;
      mu = 0.0;
      v = C{0.0, 0.0};
    }
    const double mlt = -ml / bprev;
    const double b = md + mlt * muprev;
    const C gg = v + mlt * gprev;
    a.sb[o] = b;
    sg[o] = gg;
    bprev = b;
    gprev = gg;
    muprev = mu;
  }
  C up = {0.0, 0.0};
  for (int j = ny; j >= 1; --j) {
    const long long o = (long long)j * sp + m;
    double mu;
    if (j == ny)
      mu = 0.0;
    else if (j == 1)
      mu = mean ? 0.0 : -1.0;
    else
      mu = 1.0 / (a.dy[j + 1] * a.dyf[j]);
    const C gg = sg[o];
    const double b = a.sb[o];
    const C p = (j == ny) ? C{gg.x / b, gg.y / b} : C{(gg.x - mu * up.x) / b, (gg.y - mu * up.y) / b};
    cp[o] = p;
    up = p;
  }
}

// -------------------------------------------------------------- courant ----
// Per-point time-step limits of courant (advance.f90:110-133) reduced to one
// minimum per workgroup.  Points are (i < Nx, k < Nz, j in [j0, j1]) of the
// velocity u_f:
//   spectral == 0: physical fields P[j][k][i] (the velocity just after the
//                  Crank-Nicolson solves, channel.f90:882-888);
//   spectral == 1: the spectral fields read as the real array that shares
//                  their memory (u1 and cu1 are one buffer in the reference,
//                  fft.f90:72-97), which is what courant sees when
//                  pre_first_step calls it: u(i, k, j) = Re or Im (i even or
//                  odd) of cu(i/2, k, j).  Slots k in [Nz-Nkz, Nz-1] of that
//                  buffer still hold the unnormalised z transform that the
//                  packing of fft_xz_to_fourier copied from (fft.f90:306-311),
//                  i.e. Nz times the packed value at k - Nz + 1 + 2 Nkz; the
//                  slots of the truncated modes hold that transform's
//                  round-off and are read as zero.
// dt_x = cdx / |u1 - offx|, dt_y = cdy(j) / |u2|, dt_z = cdz / |u3 + offz|
// with offx = offz = 0 except for the Front flavour (one (offx, offz) row per
// scalar).  Division by zero gives +inf, which never wins the minimum.
struct CflArgs {
  const void* u[3];
  int spectral;
  const double* cdy;   // CFL * dy(j), j = 0..Ny+1
  double cdx, cdz;     // CFL * dx, CFL * dz
  const double* offx;  // [nset][Ny+2]
  const double* offz;
  int nset;            // 1 (plain) or N_th (Front)
  int j0, j1;
  double* partial;     // one value per workgroup
};

__device__ inline double cfl_value(const CflArgs& a, const Geo& g, int f, int i, int k, int j) {
  if (!a.spectral) return reinterpret_cast<const double*>(a.u[f])[(long long)j * g.pp + (long long)k * g.nx + i];
  const int q = i >> 1;
  if (q >= g.nq) return 0.0;
  int kk = k;
  double sc = 1.0;
  if (k >= g.nkk) {
    if (k < g.nz - g.nkz) return 0.0;
    kk = k - g.nz + 1 + 2 * g.nkz;
    sc = (double)g.nz;
  }
  const double2 c = reinterpret_cast<const double2*>(a.u[f])[(long long)j * g.sp + (long long)q * g.nkk + kk];
  return sc * ((i & 1) ? c.y : c.x);
}

__global__ void __launch_bounds__(256) k_cfl_min(CflArgs a, Geo g) {
  __shared__ double red[256 / 64];
  const long long n = (long long)(a.j1 - a.j0 + 1) * g.nz * g.nx;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double best = __builtin_inf();
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int j = a.j0 + (int)(t / ((long long)g.nz * g.nx));
    const long long r = t % ((long long)g.nz * g.nx);
    const int k = (int)(r / g.nx), i = (int)(r % g.nx);
    const double u1 = cfl_value(a, g, 0, i, k, j);
    const double u2 = cfl_value(a, g, 1, i, k, j);
    const double u3 = cfl_value(a, g, 2, i, k, j);
    const double dty = a.cdy[j] / fabs(u2);
    for (int s = 0; s < a.nset; ++s) {
      const double dtx = a.cdx / fabs(u1 - a.offx[s * g.ny2 + j]);
      const double dtz = a.cdz / fabs(u3 + a.offz[s * g.ny2 + j]);
      best = fmin(best, fmin(dtx, fmin(dty, dtz)));
    }
  }
  for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = red[0];
    for (int x = 1; x < (int)(blockDim.x >> 6); ++x) v = fmin(v, red[x]);
    a.partial[blockIdx.x] = v;
  }
}

}  // namespace dgpu
