// dgpu_les.h -- subgrid-scale closures of the channel substep (for/les.f90).
//
// The reference forms the strain S_ij (compute_strain_chan, les.f90:1463-1510),
// the scaled "invariant" strain SI_ij and rotation O_ij
// (compute_strain_chan_invariant / compute_rotation_chan[_invariant],
// :1514-1675) and the velocity/scalar gradients (compute_all_gradients_chan,
// :1681-1767) in Fourier space and transforms each to physical space: 30
// inverse transforms per substep for AMD-Verstappen.  Every one of those
// tensors is a per-plane-constant linear combination of the nine velocity
// gradient components, so here the nine gradients (plus three per scalar) are
// transformed once -- the z inverse pass forms each derivative from the
// spectral state on the fly -- and the tensors are assembled pointwise in
// physical space by the model kernels below.
//
// Gradient fields (physical, plane j; GY = wall-normal velocity points):
//   G11 = d1 u1  (GYF)   G12 = d2 u1 = (u1(j)-u1(j-1))/dy(j)  (GY)   G13 = d3 u1 (GYF)
//   G21 = d1 u2  (GY)    G22 = (u2(j+1)-u2(j))/dyf(j)         (GYF)  G23 = d3 u2 (GY)
//   G31 = d1 u3  (GYF)   G32 = (u3(j)-u3(j-1))/dy(j)          (GY)   G33 = d3 u3 (GYF)
// and, per scalar, T1 = d1 th, T2 = (th(j)-th(j-1))/dy(j) (GY), T3 = d3 th.
// The reference's du2dy(j) = (u2(j)-u2(j-1))/dyf(j-1) (les.f90:1713) is G22(j-1).
#pragma once
#include "dgpu_kernels.h"
#include "les_point.h"

namespace dgpu {

// Derivative formed on the fly by the z inverse pass.  The last three are
// the wall-normal forms of poisson_p_chan (channel.f90:1085-1141): a centred
// difference (c(j+1) - c(j-1)) / (2 dyf(j)) and i k 0.5 (c(j) + c(j+1)).
enum DerivOp {
  OP_ID = 0,
  OP_DX = 1,
  OP_DZ = 2,
  OP_DY_GY = 3,
  OP_DY_GYF = 4,
  OP_DY_C2 = 5,
  OP_DX_AVG = 6,
  OP_DZ_AVG = 7
};

// Unpack + zero + backward z transform of a derivative of a spectral field
// (fft.f90:319-345 applied to the derivative arrays of les.f90).  Planes
// where the derivative is not defined (j = 0 for OP_DY_GY and OP_DY_C2,
// j = Ny+1 for OP_DY_GYF, OP_DY_C2 and the averages) are written as zero.
template <typename R, bool WAVE, int LOGN = 0>
__global__ void __launch_bounds__(512) k_zinv_op(const cpx<double>* __restrict__ C, cpx<R>* __restrict__ T,
                                                 Geo g, const cpx<R>* __restrict__ tw, int j0, int op,
                                                 const double* __restrict__ kx,
                                                 const double* __restrict__ kz,
                                                 const double* __restrict__ dy,
                                                 const double* __restrict__ dyf, double s0, double s1,
                                                 FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {  // blockIdx.z picks the field
    C = reinterpret_cast<const cpx<double>*>(fb.in[blockIdx.z]);
    T = reinterpret_cast<cpx<R>*>(fb.out[blockIdx.z]);
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nz >> 3;
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * fft_lds_len(g.nz);
  // power-formed twiddles from the caches, as in k_zinv
  const cpx<R>* twl = LOGN == FFT_XL512P ? tw : lds_twiddles(smem + ngroups * fft_lds_len(g.nz), tw, g.nz);
  const int j = j0 + blockIdx.y, q = blockIdx.x * ngroups + grp;
  bool act = q < g.nq;
  if ((op == OP_DY_GY || op == OP_DY_C2) && j == 0) act = false;
  if (op >= OP_DY_GYF && j == g.ny + 1) act = false;
  const long long base = (long long)j * g.sp + (long long)(q < g.nq ? q : 0) * g.nkk;
  const cpx<double>* col = C + base;
  const double kxq = kx[q < g.nq ? q : 0];
  double sdy = 0.0;
  if (act && op == OP_DY_GY) sdy = 1.0 / dy[j];
  if (act && op == OP_DY_GYF) sdy = 1.0 / dyf[j];
  if (act && op == OP_DY_C2) sdy = 1.0 / (2.0 * dyf[j]);
  const double sc = (q == 0) ? s0 : s1;
  const int off = g.nz - 2 * g.nkz - 1;
  cpx<R> v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int k = gt + G * r;
    cpx<double> z = {0.0, 0.0};
    int kk = -1;
    if (k <= g.nkz)
      kk = k;
    else if (k >= g.nz - g.nkz)
      kk = k - off;
    if (act && kk >= 0) {
      const cpx<double> c = col[kk];
      switch (op) {
        case OP_DX: z = ik(kxq, c); break;
        case OP_DZ: z = ik(kz[kk], c); break;
        case OP_DY_GY: z = sdy * (c - col[kk - g.sp]); break;
        case OP_DY_GYF: z = sdy * (col[kk + g.sp] - c); break;
        case OP_DY_C2: z = sdy * (col[kk + g.sp] - col[kk - g.sp]); break;
        case OP_DX_AVG: z = ik(0.5 * kxq, c + col[kk + g.sp]); break;
        case OP_DZ_AVG: z = ik(0.5 * kz[kk], c + col[kk + g.sp]); break;
        default: z = c; break;
      }
    }
    v[r] = {(R)(sc * z.x), (R)(sc * z.y)};
  }
  fft_any<R, 1, WAVE, LOGN>(v, fbuf, g.nz, g.lognz, twl, gt);
  if (q < g.nq) {
    cpx<R>* out = T + (long long)j * g.tp + (long long)q * g.nz;
    for (int k = gt; k < g.nz; k += G) st_out(out + k, fbuf[pidx(k)]);
  }
}

// ------------------------------------------------------- velocity model --
// R is the precision of the physical fields and of the model arithmetic;
// nu_t and kappa_t are stored in FP64 for the CN solve.
template <typename R>
struct LesVelArgs {
  const R* G[9];        // G11 G12 G13 G21 G22 G23 G31 G32 G33 (see above)
  const R* u1;
  const R* u3;
  R* tau11;             // out: delta_yf * s1 * S11 etc. (GYF), delta_y * temp * (S12 - ...) (GY)
  R* tau33;
  R* tau13;
  R* tau12;
  R* tau23;
  R* cr1;               // out: nu_t (u1(j) - u1(j-1))/dy(j)  (les.f90:893)
  R* cr3;               // out: nu_t (u3(j) - u3(j-1))/dy(j)  (les.f90:908)
  double* nu_t;         // out
  const double* dy;
  const double* dyf;
  double dx1, dz1;      // Lx/Nx, Lz/Nz
  int type;             // 1 Smagorinsky, 4 AMD-Rozema, 5 AMD-Verstappen
  int js, je;           // jstart, jend
};

using lespt::kBetaSgs;
using lespt::kCAmd;
using lespt::kCSmag;

// Per-plane constants of the velocity model at plane j (les.f90:146-222,
// 371-518, 593-799): which of the GYF / GY evaluations the reference makes
// there, and the filter factors, formed in FP64 from the FP64 metrics.
template <typename R>
__device__ inline lespt::VelPlane<R> les_vel_plane(const LesVelArgs<R>& a, const Geo& g, int j) {
  const int ny = g.ny;
  lespt::VelPlane<R> P;
  P.type = a.type;
  const bool gyf_rows = (a.type == 5) ? (j >= a.js && j <= a.je) : (j <= ny);  // :172-182, 468-478
  P.gyf_in = gyf_rows && j <= ny;  // the GYF model also reads GY plane j+1
  if (a.type == 1)
    P.gy_in = true;  // :146-168, 185-197: planes 1..Ny+1
  else if (a.type == 4)
    P.gy_in = (j >= 2);  // :371-464, 482-494
  else
    P.gy_in = (j >= 2 && j <= ny);  // :647-738 (Ny+1 never consumed)
  P.nt_in = P.gy_in && j >= 3 && j <= ny - 1;  // ghost_les_mpi zeroes nu_t at 1, 2, Ny, Ny+1
  P.cr_in = (j >= 2 && j <= ny);
  P.dyj = (R)a.dy[j];
  P.dyfj = (R)a.dyf[j];
  P.dyfm = (R)a.dyf[j - 1];
  P.dysu = (j <= ny) ? (R)(a.dy[j + 1] * 2.0) : (R)0;
  const double ddxs = a.dx1 * kBetaSgs, ddzs = a.dz1 * kBetaSgs, ddy = a.dy[j], ddyf = a.dyf[j];
  P.dxs = (R)ddxs;
  P.dzs = (R)ddzs;
  double dyf_f, dy_f;  // les.f90:203-222, 499-518, 775-799
  if (a.type == 1) {
    const double c2 = -2.0 * kCSmag * kCSmag;
    dyf_f = c2 * pow(ddxs * ddyf * 1.0 * ddzs, 2.0 / 3.0);
    dy_f = c2 * pow(ddxs * ddy * 2.0 * ddzs, 2.0 / 3.0);
  } else if (a.type == 4) {
    dyf_f = dy_f = -2.0 * kCAmd * kCAmd;
  } else {
    const double c2 = -2.0 * kCAmd * kCAmd * 3.0;
    dyf_f = c2 / (1.0 / (ddxs * ddxs) + 1.0 / ((ddyf * 2.0) * (ddyf * 2.0)) + 1.0 / (ddzs * ddzs));
    dy_f = c2 / (1.0 / (ddxs * ddxs) + 1.0 / ((ddy * 2.0) * (ddy * 2.0)) + 1.0 / (ddzs * ddzs));
  }
  P.dyf_fac = (R)dyf_f;
  P.dy_fac = (R)dy_f;
  return P;
}

// The model's inputs at point o of plane j (plane j+1 only where it exists).
template <typename R>
__device__ inline void les_vel_load(const LesVelArgs<R>& a, const Geo& g, int j, long long o,
                                    lespt::VelIn<R>& x) {
  const long long pp = g.pp;
  for (int m = 0; m < 9; ++m) x.c[m] = a.G[m][o];
  const int upm[4] = {1, 3, 5, 7};   // G12 G21 G23 G32
  const int dnm[5] = {0, 4, 8, 2, 6};  // G11 G22 G33 G13 G31
  for (int t = 0; t < 4; ++t) x.up[t] = (j <= g.ny) ? a.G[upm[t]][o + pp] : (R)0;
  for (int t = 0; t < 5; ++t) x.dn[t] = a.G[dnm[t]][o - pp];
  x.u1o = a.u1[o];
  x.u1m = a.u1[o - pp];
  x.u3o = a.u3[o];
  x.u3m = a.u3[o - pp];
}

// --------------------------------------------------------- scalar model --
template <typename R>
struct LesThArgs {
  const R* G[9];
  const R* T1;        // d1 th (GYF)
  const R* T2;        // (th(j)-th(j-1))/dy(j) (GY)
  const R* T3;        // d3 th (GYF)
  const double* nu_t;
  R* f1;              // out: flux to be x-differentiated (GYF)
  R* f3;              // out: flux to be z-differentiated (GYF)
  double* kappa_t;    // out
  const double* dy;
  const double* dyf;
  double dx1, dz1;
  int type;
  int jst, jet;
};

// Per-plane constants of the AMD scalar model at plane j (les.f90:1238-1384).
template <typename R>
__device__ inline lespt::ThPlane<R> les_th_plane(const LesThArgs<R>& a, const Geo& g, int j) {
  const int ny = g.ny;
  lespt::ThPlane<R> P;
  P.gyf_in = (j >= a.jst && j <= a.jet && j <= ny);
  P.gy_in = (j >= 3 && j <= ny - 1);
  const double ddxs = a.dx1 * kBetaSgs, ddzs = a.dz1 * kBetaSgs;
  P.dxs = (R)ddxs;
  P.dzs = (R)ddzs;
  P.dyj = (R)a.dy[j];
  P.dyfj = (R)a.dyf[j];
  P.dyfm = (R)a.dyf[j - 1];
  const double c2 = -2.0 * kCAmd * kCAmd * 3.0;
  const double ddlf = a.dyf[j] * 2.0, ddl = a.dy[j] * 2.0;
  P.dyf_fac = (R)(c2 / (1.0 / (ddxs * ddxs) + 1.0 / (ddlf * ddlf) + 1.0 / (ddzs * ddzs)));
  P.dy_fac = (R)(c2 / (1.0 / (ddxs * ddxs) + 1.0 / (ddl * ddl) + 1.0 / (ddzs * ddzs)));
  return P;
}

template <typename R>
__device__ inline void les_th_load(const LesThArgs<R>& a, const Geo& g, int j, long long o, lespt::ThIn<R>& x) {
  const long long pp = g.pp;
  for (int m = 0; m < 9; ++m) x.c[m] = a.G[m][o];
  x.t[0] = a.T1[o];
  x.t[1] = a.T2[o];
  x.t[2] = a.T3[o];
  const bool up = j <= g.ny;
  x.up[0] = up ? a.G[1][o + pp] : (R)0;
  x.up[1] = up ? a.G[3][o + pp] : (R)0;
  x.up[2] = up ? a.G[5][o + pp] : (R)0;
  x.up[3] = up ? a.G[7][o + pp] : (R)0;
  x.up[4] = up ? a.T2[o + pp] : (R)0;
  x.dn[0] = a.G[0][o - pp];
  x.dn[1] = a.G[4][o - pp];
  x.dn[2] = a.G[8][o - pp];
  x.dn[3] = a.G[2][o - pp];
  x.dn[4] = a.G[6][o - pp];
  x.dn[5] = a.T1[o - pp];
  x.dn[6] = a.T3[o - pp];
  x.dd = (j >= 2) ? a.G[4][o - 2 * pp] : (R)0;
}

// One thread per physical point of planes 1..Ny+1, both subgrid models.
// Velocity: the eddy viscosity, the modelled stresses scaled by the filter
// factors of les.f90:201-222, 498-518, 773-799 (applied here before the
// forward transform instead of after it), and the explicit cross-term fluxes
// (les.f90:890-919); the point model is lespt::vel_point.  Values at points
// the substep never consumes (GY plane 1, GYF plane Ny+1, the zeroed nu_t
// planes) follow the reference where they are cheap and are otherwise left
// as the raw strain.  Scalars, each: types 1/4 kappa_t = nu_t and the
// explicit horizontal diffusion fluxes kappa_t d1 th, kappa_t d3 th
// (les.f90:1142-1203); type 5 the AMD scalar model at GYF (:1238-1275) and
// GY points (:1278-1330), kappa_t (:1373-1384) and the fluxes
// 0.5 delta_yf s1_th d_i th (:1334-1342, 1392-1401), with the filter factor
// applied before the forward transform (lespt::th_point).  The scalar model
// reads the same velocity gradients (planes and components) as the velocity
// model: they are taken from registers instead of being read again.
template <typename R>
struct LesThAll {
  LesThArgs<R> s[MAXTH];
  int nth;
};

// The four y-difference gradients -- G12 = (u1(j) - u1(j-1))/dy(j),
// G32 likewise, G22 = (u2(j+1) - u2(j))/dyf(j) (0 at Ny+1) and per scalar
// T2 = (th(j) - th(j-1))/dy(j) -- commute with the x-z transforms, so they are
// formed here from the physical state instead of by four inverse transforms
// of the differenced spectra (the forms of k_zinv_op's OP_DY_GY / OP_DY_GYF).
// With store set (the adjoint's recomputation) they are also written out
// (planes 0..Ny+1), where the adjoint's gathers read them.
template <typename R>
struct LesDyArgs {
  const R* u2;
  const R* th[MAXTH];
  R* g12;
  R* g22;
  R* g32;
  R* t2[MAXTH];
  int store;
};

template <typename R>
__global__ void __launch_bounds__(256) k_les_model(LesVelArgs<R> a, LesThAll<R> th, LesDyArgs<R> dy, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int ny = g.ny;
  const int j = 1 + plane_of(idx, pp);
  if (j > ny + 1) return;
  const long long o = (long long)j * pp + (long long)((unsigned)idx % (unsigned)pp);
  const bool up = j <= ny;  // plane j+1 exists
  const R idy = (R)(1.0 / a.dy[j]), idyp = up ? (R)(1.0 / a.dy[j + 1]) : (R)0;
  // y differences at planes j (and j+1, j-1, j-2 where the models read them)
  const R u1m = a.u1[o - pp], u1c = a.u1[o], u1p = up ? a.u1[o + pp] : (R)0;
  const R u3m = a.u3[o - pp], u3c = a.u3[o], u3p = up ? a.u3[o + pp] : (R)0;
  const R u2m = dy.u2[o - pp], u2c = dy.u2[o], u2p = up ? dy.u2[o + pp] : (R)0;
  const R g12 = (u1c - u1m) * idy, g12p = up ? (u1p - u1c) * idyp : (R)0;
  const R g32 = (u3c - u3m) * idy, g32p = up ? (u3p - u3c) * idyp : (R)0;
  const R g22 = up ? (u2p - u2c) * (R)(1.0 / a.dyf[j]) : (R)0;
  const R g22m = (u2c - u2m) * (R)(1.0 / a.dyf[j - 1]);
  const lespt::VelPlane<R> P = les_vel_plane(a, g, j);
  lespt::VelIn<R> x;
  for (int m = 0; m < 9; ++m) x.c[m] = (m == 1 || m == 4 || m == 7) ? (R)0 : a.G[m][o];
  x.c[1] = g12;
  x.c[4] = g22;
  x.c[7] = g32;
  x.up[0] = g12p;
  x.up[1] = up ? a.G[3][o + pp] : (R)0;
  x.up[2] = up ? a.G[5][o + pp] : (R)0;
  x.up[3] = g32p;
  x.dn[0] = a.G[0][o - pp];
  x.dn[1] = g22m;
  x.dn[2] = a.G[8][o - pp];
  x.dn[3] = a.G[2][o - pp];
  x.dn[4] = a.G[6][o - pp];
  x.u1o = u1c;
  x.u1m = u1m;
  x.u3o = u3c;
  x.u3m = u3m;
  lespt::VelOut<R> y;
  lespt::vel_point(P, x, y);
  a.tau11[o] = y.t11;
  a.tau33[o] = y.t33;
  a.tau13[o] = y.t13;
  a.tau12[o] = y.t12;
  a.tau23[o] = y.t23;
  a.nu_t[o] = (double)y.nt;
  a.cr1[o] = y.cr1;
  a.cr3[o] = y.cr3;
  if (dy.store) {
    dy.g12[o] = g12;
    dy.g22[o] = g22;
    dy.g32[o] = g32;
    if (j == 1) {  // plane 0: the differences the transform route gives there
      dy.g12[o - pp] = (R)0;
      dy.g32[o - pp] = (R)0;
      dy.g22[o - pp] = g22m;
    }
  }
  for (int n = 0; n < th.nth; ++n) {
    const LesThArgs<R>& t = th.s[n];
    const R thm = dy.th[n][o - pp], thc = dy.th[n][o], thp = up ? dy.th[n][o + pp] : (R)0;
    const R t2 = (thc - thm) * idy, t2p = up ? (thp - thc) * idyp : (R)0;
    if (dy.store) {
      dy.t2[n][o] = t2;
      if (j == 1) dy.t2[n][o - pp] = (R)0;
    }
    if (t.type != 5) {
      const double kt = (double)y.nt;
      t.kappa_t[o] = kt;
      t.f1[o] = (R)kt * t.T1[o];
      t.f3[o] = (R)kt * t.T3[o];
      continue;
    }
    const lespt::ThPlane<R> Pt = les_th_plane(t, g, j);
    lespt::ThIn<R> xt;
#pragma unroll
    for (int m = 0; m < 9; ++m) xt.c[m] = x.c[m];
    xt.t[0] = t.T1[o];
    xt.t[1] = t2;
    xt.t[2] = t.T3[o];
#pragma unroll
    for (int m = 0; m < 4; ++m) xt.up[m] = x.up[m];  // G12 G21 G23 G32 at j+1
    xt.up[4] = t2p;
#pragma unroll
    for (int m = 0; m < 5; ++m) xt.dn[m] = x.dn[m];  // G11 G22 G33 G13 G31 at j-1
    xt.dn[5] = t.T1[o - pp];
    xt.dn[6] = t.T3[o - pp];
    xt.dd = (j >= 2) ? (dy.u2[o - pp] - dy.u2[o - 2 * pp]) * (R)(1.0 / a.dyf[j - 2]) : (R)0;
    lespt::ThOut<R> yt;
    lespt::th_point(Pt, xt, yt);
    t.f1[o] = yt.f1;
    t.f3[o] = yt.f3;
    t.kappa_t[o] = (double)yt.kt;
  }
}

// ---------------------------------------------- stress divergence into cf --
// les.f90:861-919 (velocity) and :1173-1203 / :1404-1413 (scalar), written as
// separate forcing arrays that the spectral RHS pass adds to cf.
struct LesDivArgs {
  const double2* t11;
  const double2* t33;
  const double2* t13;
  const double2* t12;
  const double2* t23;
  const double2* c1;
  const double2* c3;
  double2* f1;
  double2* f2;
  double2* f3;
  const double2* th_f1[MAXTH];
  const double2* th_f3[MAXTH];
  double2* fth[MAXTH];
  const double* kx;
  const double* kz;
  const double* dyf;
  int js, j2;         // cf1/cf3 rows [jstart, Ny-1] (les.f90:63-66, 861)
  int nth;
  int th_sign;        // +1 for types 1/4 (les.f90:1176, 1200), -1 for type 5 (:1407-1410)
  int jst[MAXTH], jet[MAXTH];
};

__global__ void __launch_bounds__(256) k_les_div(LesDivArgs a, Geo g) {
  typedef cpx<double> C;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long sp = g.sp;
  const int j = plane_of(idx, sp);
  if (j >= g.ny2) return;
  const long long m = idx % sp;
  const int q = (int)((unsigned)m / (unsigned)g.nkk), kk = (int)((unsigned)m % (unsigned)g.nkk);
  const double kxq = a.kx[q], kzk = a.kz[kk];
  const long long o = idx;
  auto L = [&](const double2* p, long long i) { return reinterpret_cast<const C*>(p)[i]; };
  C f1 = {0, 0}, f2 = {0, 0}, f3 = {0, 0};
  if (j >= a.js && j <= a.j2) {
    const double r = 1.0 / a.dyf[j];
    f1 = C{0, 0} - ik(kxq, L(a.t11, o)) - r * (L(a.t12, o + sp) - L(a.t12, o)) - ik(kzk, L(a.t13, o));
    f3 = C{0, 0} - ik(kxq, L(a.t13, o)) - r * (L(a.t23, o + sp) - L(a.t23, o)) - ik(kzk, L(a.t33, o));
  }
  if (j >= 2 && j <= g.ny) {
    f2 = C{0, 0} - ik(kxq, L(a.t12, o)) - ik(kzk, L(a.t23, o));
    f2 = f2 + ik(kxq, L(a.c1, o)) + ik(kzk, L(a.c3, o));
  }
  reinterpret_cast<C*>(a.f1)[o] = f1;
  reinterpret_cast<C*>(a.f2)[o] = f2;
  reinterpret_cast<C*>(a.f3)[o] = f3;
  for (int n = 0; n < a.nth; ++n) {
    C ft = {0, 0};
    if (j >= a.jst[n] && j <= a.jet[n])
      ft = (double)a.th_sign * (ik(kxq, L(a.th_f1[n], o)) + ik(kzk, L(a.th_f3[n], o)));
    reinterpret_cast<C*>(a.fth[n])[o] = ft;
  }
}

// -------------------------------------------------------- LES BCs --------
// Scalar post BCs (boundary.f90:728-821) and the Neumann ghost copies of
// apply_BC_les / apply_BC_th_les (les.f90:2001-2100) on the spectral state.
struct LesBcArgs {
  double2* cu1;
  double2* cu3;
  double2* cth[MAXTH];
  const double* dy;
  int nth;
  int u_lo, u_hi, w_lo, w_hi;
  int th_lo[MAXTH], th_hi[MAXTH];
  double th_lo_c1[MAXTH], th_hi_c1[MAXTH];
};

__global__ void k_les_bc(LesBcArgs a, Geo g) {
  typedef cpx<double> C;
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.sp) return;
  const bool mean = (m == 0);
  const long long P = g.sp;
  const int ny = g.ny;
  const long long T0 = (long long)ny * P + m, T1 = T0 + P, Tm = T0 - P;
  for (int n = 0; n < a.nth; ++n) {
    C* c = reinterpret_cast<C*>(a.cth[n]);
    if (a.th_lo[n] == 0) {
      c[m] = C{0, 0};
      c[P + m] = C{0, 0};
      if (mean) c[P + m] = c[m] = C{a.th_lo_c1[n], 0.0};
    } else {
      c[P + m] = c[2 * P + m];
      c[m] = c[P + m];
      if (mean) {
        c[P + m].x = c[2 * P + m].x - a.dy[2] * a.th_lo_c1[n];
        c[m] = C{c[P + m].x - a.dy[1] * a.th_lo_c1[n], c[P + m].y};
      }
    }
    if (a.th_hi[n] == 0) {
      c[T0] = C{0, 0};
      c[T1] = C{0, 0};
      if (mean) c[T0] = c[T1] = C{a.th_hi_c1[n], 0.0};
    } else {
      c[T0] = c[Tm];
      c[T1] = c[T0];
      if (mean) {
        c[T0].x = c[Tm].x + a.dy[ny] * a.th_hi_c1[n];
        c[T1] = C{c[T0].x + a.dy[ny] * a.th_hi_c1[n], c[T0].y};
      }
    }
    // apply_BC_th_les
    if (a.th_hi[n] == 1) {
      c[T0] = c[Tm];
      c[T1] = c[T0];
    }
    if (a.th_lo[n] == 1) {
      c[P + m] = c[2 * P + m];
      c[m] = c[P + m];
    }
  }
  C* u1 = reinterpret_cast<C*>(a.cu1);
  C* u3 = reinterpret_cast<C*>(a.cu3);
  if (a.u_hi == 1) {
    u1[T0] = u1[Tm];
    u1[T1] = u1[T0];
  }
  if (a.w_hi == 1) {
    u3[T0] = u3[Tm];
    u3[T1] = u3[T0];
  }
  if (a.u_lo == 1) {
    u1[P + m] = u1[2 * P + m];
    u1[m] = u1[P + m];
  }
  if (a.w_lo == 1) {
    u3[P + m] = u3[2 * P + m];
    u3[m] = u3[P + m];
  }
}

// ------------------------------------------------- CN solve with nu_t ----
// The real CN solve when the rows depend on the eddy viscosity/diffusivity
// (channel.f90:676-689, 736-749, 792-805, 848-861, with the explicit LES terms
// of :519-547, 634-646): one (x,z) column per thread, the Thomas sweeps of
// :1335-1496 run sequentially, through two scratch arrays (sb, sg).
struct TriRow {
  double a, b, c;   // lane-independent part of the row (all of it for fixed rows)
  double vb;        // fixed-row right-hand side (scaled by kscale when kind == 2)
  double wp, wm;    // explicit stencil weights (assembled rows)
  double lp, lm;    // LES weights: temp2/(dy(j+1) dyf(j)), temp2/(dy(j) dyf(j)) (u type) or
                    // temp2/(dyf(j) dy(j)), temp2/(dyf(j-1) dy(j)) (v type)
  int kind;         // 1 assembled, 0 fixed, 2 fixed scaled
  int vtype;        // 1 for u2 (GY rows: nu_t averaged to GYF)
};

// T is the storage type of r and u; the sweeps run in FP64 with their
// scratch in FP64.  incr has the meaning of k_thomas_seg's flag: the solve is for
// d = x - u, with the right-hand side vec - A u formed in the forward sweep
// from the u(j-1), u(j), u(j+1) it holds anyway (reading u back in the
// backward sweep would put a load behind every store of the serial sweep).
// Each lane solves C adjacent columns (one C-wide access per row and array),
// so a wavefront's row accesses are 64 C values long.
//
// LesCnFields batches the solves of several fields into one launch
// (blockIdx.y = field, each with its own two scratch arrays of n pp values; nf = 0:
// the single field of the plain arguments).  A solve is one sequential sweep
// per lane, so one field fills only a few waves per CU; the batch gives the
// CUs the other fields' sweeps to hide their latency behind.
template <typename T>
struct LesCnFields {
  const T* r[3 + MAXTH];
  T* u[3 + MAXTH];
  const double* nt[3 + MAXTH];
  const TriRow* rows[3 + MAXTH];
  int ca[3 + MAXTH], cb[3 + MAXTH];
  int nf;
};

template <typename T, int C>
__global__ void __launch_bounds__(256) k_thomas_les(const T* __restrict__ r, T* __restrict__ u,
                                                    const double* __restrict__ nt, double* __restrict__ sb,
                                                    double* __restrict__ sg, const TriRow* __restrict__ rows,
                                                    int n, long long pp, int ca, int cb, double kscale,
                                                    int incr, LesCnFields<T> mf = LesCnFields<T>{}) {
  if (mf.nf > 0) {
    const int f = blockIdx.y;
    r = mf.r[f];
    u = mf.u[f];
    nt = mf.nt[f];
    rows = mf.rows[f];
    ca = mf.ca[f];
    cb = mf.cb[f];
    sb += (long long)f * n * pp;
    sg += (long long)f * n * pp;
  }
  typedef T VT __attribute__((ext_vector_type(C)));
  typedef double VD __attribute__((ext_vector_type(C)));
  const long long col = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * C;
  if (col >= pp) return;
  auto at = [&](int j) { return (long long)j * pp + col; };
  auto U = [&](int j) { return __builtin_convertvector(*reinterpret_cast<const VT*>(u + at(j)), VD); };
  auto N = [&](int j) { return *reinterpret_cast<const VD*>(nt + at(j)); };
  auto R = [&](int j) { return __builtin_convertvector(*reinterpret_cast<const VT*>(r + at(j)), VD); };
  const VD zero = (VD)0.0;
  // row coefficients of this lane's columns
  auto row = [&](int j, VD& a, VD& b, VD& c, VD& vec, VD um, VD uc, VD up) {
    const TriRow& rw = rows[j];
    a = (VD)rw.a;
    b = (VD)rw.b;
    c = (VD)rw.c;
    if (rw.kind == 1) {
      VD nl, nu;  // eddy viscosity on the lower / upper face of the row
      if (rw.vtype) {
        nl = 0.5 * (N(j) + N(j - 1));
        nu = 0.5 * (N(j) + N(j + 1));
      } else {
        nl = N(j);
        nu = N(j + 1);
      }
      a -= rw.lm * nl;
      b += rw.lp * nu + rw.lm * nl;
      c -= rw.lp * nu;
      vec = R(j) + (incr ? uc : zero) + rw.wp * (up - uc) - rw.wm * (uc - um) +
            (rw.lp * nu * (up - uc) - rw.lm * nl * (uc - um));
    } else {
      vec = (VD)((rw.kind == 2) ? rw.vb * kscale : rw.vb);
    }
    if (incr) vec -= a * um + b * uc + c * up;
  };
  // The forward sweep stores e(j) = g'(j)/b'(j) and f(j) = c(j)/b'(j) (c with
  // the eddy viscosity of the row's upper face), so that the backward sweep
  // x(j) = e(j) - f(j) x(j+1) reads two scratch values a row and neither the
  // eddy viscosity again nor divides: 64 instead of 72 bytes a point in FP64.
  VD um = zero, uc = U(0), up = (n > 1) ? U(1) : zero;
  VD a, b, c, vec;
  row(0, a, b, c, vec, um, uc, up);
  VD inv = 1.0 / b, gp = vec;
  *reinterpret_cast<VD*>(sb + at(0)) = c * inv;
  *reinterpret_cast<VD*>(sg + at(0)) = gp * inv;
  for (int j = 1; j < n; ++j) {
    um = uc;
    uc = up;
    up = (j + 1 < n) ? U(j + 1) : zero;
    const VD cprev = c;
    row(j, a, b, c, vec, um, uc, up);
    const VD mlt = -a * inv;  // -a(j)/b'(j-1)
    const VD bp = b + mlt * cprev;
    gp = vec + mlt * gp;
    inv = 1.0 / bp;
    *reinterpret_cast<VD*>(sb + at(j)) = c * inv;
    *reinterpret_cast<VD*>(sg + at(j)) = gp * inv;
  }
  // backward: x(n-1) = g'/b' = e(n-1), x(j) = e(j) - f(j) x(j+1)
  VD x = gp * inv;
  auto put = [&](int j, VD v) { *reinterpret_cast<VT*>(u + at(j)) = __builtin_convertvector(v, VT); };
  if (n - 1 >= ca && n - 1 <= cb) put(n - 1, x);
  for (int j = n - 2; j >= ca; --j) {
    x = *reinterpret_cast<const VD*>(sg + at(j)) - *reinterpret_cast<const VD*>(sb + at(j)) * x;
    if (j <= cb) put(j, x);
  }
}

}  // namespace dgpu
