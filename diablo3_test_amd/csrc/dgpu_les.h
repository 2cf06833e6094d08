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
template <typename R, bool WAVE>
__global__ void __launch_bounds__(512) k_zinv_op(const cpx<double>* __restrict__ C, cpx<R>* __restrict__ T,
                                                 Geo g, const cpx<R>* __restrict__ tw, int j0, int op,
                                                 const double* __restrict__ kx,
                                                 const double* __restrict__ kz,
                                                 const double* __restrict__ dy,
                                                 const double* __restrict__ dyf, double s0, double s1) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nz >> 3;
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * fft_lds_len(g.nz);
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
  fft_regs<R, 1, WAVE>(v, fbuf, g.nz, g.lognz, tw, gt);
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

constexpr double kBetaSgs = 3.0;   // les.f90:60
constexpr double kCSmag = 0.17;    // les.f90:41
constexpr double kCAmd = 0.2887;   // les.f90:42

// GYF <-> GY interpolation weights of les.f90:150-165 and 316-327.
template <typename R>
__device__ __forceinline__ R gy_of(const R* a, long long o, long long pp, R dyfm, R dyfj, R dyj) {
  return (a[o] * dyfm + a[o - pp] * dyfj) / ((R)2 * dyj);
}

// -(I3 - I4) / (I1 - I2) clipped at 0 for a strain s and rotation w (AMD
// numerator of les.f90:597-637 with unit filter weights, or :311-361 with
// the Rozema weights dxq, dyq, dzq).
template <typename R>
__device__ __forceinline__ R amd_value(R a1, R a2, R a3, R a4, R a5, R a6, R w4, R w5, R w6, R dxq,
                                       R dyq, R dzq, bool rozema) {
  R v;
  if (rozema) {
    v = dxq * a1 * a1 * a1 + dyq * a2 * a2 * a2 + dzq * a3 * a3 * a3 + (2 * dxq + dyq) * a1 * a4 * a4 +
        (2 * dxq + dzq) * a1 * a5 * a5 + (2 * dyq + dxq) * a2 * a4 * a4 + (2 * dyq + dzq) * a2 * a6 * a6 +
        (2 * dzq + dxq) * a3 * a5 * a5 + (2 * dzq + dyq) * a3 * a6 * a6 +
        2 * (dxq + dyq + dzq) * a4 * a5 * a6;
    v = -v - dyq * a1 * w4 * w4 - dzq * a1 * w5 * w5 - dxq * a2 * w4 * w4 - dzq * a2 * w6 * w6 -
        dxq * a3 * w5 * w5 - dyq * a3 * w6 * w6 - 2 * dzq * a4 * w5 * w6 + 2 * dyq * a5 * w4 * w6 -
        2 * dxq * a6 * w4 * w5;
  } else {
    v = a1 * a1 * a1 + a2 * a2 * a2 + a3 * a3 * a3 + 3 * a1 * a4 * a4 + 3 * a1 * a5 * a5 + 3 * a2 * a4 * a4 +
        3 * a2 * a6 * a6 + 3 * a3 * a5 * a5 + 3 * a3 * a6 * a6 + 6 * a4 * a5 * a6;
    v = -v - a1 * w4 * w4 - a1 * w5 * w5 - a2 * w4 * w4 - a2 * w6 * w6 - a3 * w5 * w5 - a3 * w6 * w6 -
        2 * a4 * w5 * w6 + 2 * a5 * w4 * w6 - 2 * a6 * w4 * w5;
  }
  if (v <= (R)0) return (R)0;
  return v / (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 +
              2 * a6 * a6 + 2 * w6 * w6);
}

// One thread per physical point of planes 1..Ny+1: the eddy viscosity, the
// modelled stresses scaled by the filter factors of les.f90:201-222, 498-518,
// 773-799 (applied here before the forward transform instead of after it),
// and the explicit cross-term fluxes.  Values at points the substep never
// consumes (GY plane 1, GYF plane Ny+1, the zeroed nu_t planes) follow the
// reference where they are cheap and are otherwise left as the raw strain.
template <typename R>
__global__ void __launch_bounds__(256) k_les_vel(LesVelArgs<R> a, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int j = 1 + (int)(idx / pp);
  if (j > g.ny + 1) return;
  const long long o = (long long)j * pp + idx % pp;
  const int ny = g.ny;
  const R h5 = (R)0.5, two = (R)2, four = (R)4, one = (R)1, zero = (R)0;
  const R* G11 = a.G[0];
  const R* G12 = a.G[1];
  const R* G13 = a.G[2];
  const R* G21 = a.G[3];
  const R* G22 = a.G[4];
  const R* G23 = a.G[5];
  const R* G31 = a.G[6];
  const R* G32 = a.G[7];
  const R* G33 = a.G[8];
  const R dyj = (R)a.dy[j], dyfj = (R)a.dyf[j], dyfm = (R)a.dyf[j - 1];
  const R dxs = (R)(a.dx1 * kBetaSgs), dzs = (R)(a.dz1 * kBetaSgs);
  // strain at this plane: GYF components S11 S22 S33 S13, GY components S12 S23
  const R S11 = G11[o], S22 = G22[o], S33 = G33[o];
  const R S13 = h5 * (G13[o] + G31[o]);
  const R S12 = h5 * (G12[o] + G21[o]);
  const R S23 = h5 * (G23[o] + G32[o]);
  // ---- GYF point (plane j)
  R s1 = zero;
  bool gyf_in;
  if (a.type == 5)
    gyf_in = (j >= a.js && j <= a.je);
  else
    gyf_in = (j <= ny);  // types 1, 4 scale planes 1..Ny (les.f90:172-182, 468-478)
  const bool have_up = (j <= ny);  // GY plane j+1 exists
  if (gyf_in && have_up) {
    const long long u = o + pp;
    if (a.type == 1) {
      const R s12 = h5 * (h5 * (G12[u] + G21[u]) + S12);
      const R s23 = h5 * (h5 * (G23[u] + G32[u]) + S23);
      s1 = sqrt(two * S11 * S11 + four * s12 * s12 + four * S13 * S13 + two * S22 * S22 + four * s23 * s23 +
                two * S33 * S33);
    } else if (a.type == 4) {
      const R dxq = dxs * dxs, dzq = dzs * dzs, dyq = (dyj * two) * (dyj * two);
      const R a4 = h5 * (h5 * (G12[u] + G21[u]) + S12);
      const R a6 = h5 * (h5 * (G23[u] + G32[u]) + S23);
      const R w4 = h5 * (h5 * (G12[u] - G21[u]) + h5 * (G12[o] - G21[o]));
      const R w5 = h5 * (G13[o] - G31[o]);
      const R w6 = h5 * (h5 * (G23[u] - G32[u]) + h5 * (G23[o] - G32[o]));
      s1 = amd_value(S11, S22, S33, a4, S13, a6, w4, w5, w6, dxq, dyq, dzq, true);
    } else {
      // invariant strain/rotation: GY components scaled with dys = 2 dy at
      // their own plane, GYF ones with dxs, dzs (les.f90:1527-1559, 1633-1662)
      const R dysu = (R)a.dy[j + 1] * two, dysj = dyj * two;
      const R si4u = h5 * (G12[u] * (dysu / dxs) + G21[u] * (dxs / dysu));
      const R si4j = h5 * (G12[o] * (dysj / dxs) + G21[o] * (dxs / dysj));
      const R si6u = h5 * (G23[u] * (dzs / dysu) + G32[u] * (dysu / dzs));
      const R si6j = h5 * (G23[o] * (dzs / dysj) + G32[o] * (dysj / dzs));
      const R o4u = h5 * (G12[u] * (dysu / dxs) - G21[u] * (dxs / dysu));
      const R o4j = h5 * (G12[o] * (dysj / dxs) - G21[o] * (dxs / dysj));
      const R o6u = h5 * (G23[u] * (dzs / dysu) - G32[u] * (dysu / dzs));
      const R o6j = h5 * (G23[o] * (dzs / dysj) - G32[o] * (dysj / dzs));
      const R si5 = h5 * (G13[o] * (dzs / dxs) + G31[o] * (dxs / dzs));
      const R o5 = h5 * (G13[o] * (dzs / dxs) - G31[o] * (dxs / dzs));
      s1 = amd_value(S11, S22, S33, h5 * (si4u + si4j), si5, h5 * (si6u + si6j), h5 * (o4u + o4j), o5,
                     h5 * (o6u + o6j), one, one, one, false);
    }
  }
  // ---- GY point (plane j)
  R temp = zero;
  bool gy_in;
  if (a.type == 1)
    gy_in = true;                      // les.f90:146-168, 185-197: planes 1..Ny+1
  else if (a.type == 4)
    gy_in = (j >= 2);                  // :371-464, 482-494: planes 2..Ny+1
  else
    gy_in = (j >= 2 && j <= ny);       // :647-738 computes 2..Ny (Ny+1 never consumed)
  if (gy_in) {
    const long long m = o - pp;
    const R i11 = gy_of(G11, o, pp, dyfm, dyfj, dyj);
    const R i22 = gy_of(G22, o, pp, dyfm, dyfj, dyj);
    const R i33 = gy_of(G33, o, pp, dyfm, dyfj, dyj);
    if (a.type == 1) {
      const R i13 = h5 * (gy_of(G13, o, pp, dyfm, dyfj, dyj) + gy_of(G31, o, pp, dyfm, dyfj, dyj));
      temp = sqrt(two * i11 * i11 + four * S12 * S12 + four * i13 * i13 + two * i22 * i22 + four * S23 * S23 +
                  two * i33 * i33);
    } else if (a.type == 4) {
      const R dxq = dxs * dxs, dzq = dzs * dzs, dyq = (dyj * two) * (dyj * two);
      const R i13 = h5 * (gy_of(G13, o, pp, dyfm, dyfj, dyj) + gy_of(G31, o, pp, dyfm, dyfj, dyj));
      const R w4 = h5 * (G12[o] - G21[o]);
      const R w5 = (h5 * (G13[o] - G31[o]) * dyfm + h5 * (G13[m] - G31[m]) * dyfj) / (two * dyj);
      const R w6 = h5 * (G23[o] - G32[o]);
      temp = amd_value(i11, i22, i33, S12, i13, S23, w4, w5, w6, dxq, dyq, dzq, true);
    } else {
      const R dysj = dyj * two;
      const R si4 = h5 * (G12[o] * (dysj / dxs) + G21[o] * (dxs / dysj));
      const R si6 = h5 * (G23[o] * (dzs / dysj) + G32[o] * (dysj / dzs));
      const R o4 = h5 * (G12[o] * (dysj / dxs) - G21[o] * (dxs / dysj));
      const R o6 = h5 * (G23[o] * (dzs / dysj) - G32[o] * (dysj / dzs));
      const R si5o = h5 * (G13[o] * (dzs / dxs) + G31[o] * (dxs / dzs));
      const R si5m = h5 * (G13[m] * (dzs / dxs) + G31[m] * (dxs / dzs));
      const R o5o = h5 * (G13[o] * (dzs / dxs) - G31[o] * (dxs / dzs));
      const R o5m = h5 * (G13[m] * (dzs / dxs) - G31[m] * (dxs / dzs));
      const R si5 = (si5o * dyfm + si5m * dyfj) / (two * dyj);
      const R o5 = (o5o * dyfm + o5m * dyfj) / (two * dyj);
      temp = amd_value(i11, i22, i33, si4, si5, si6, o4, o5, o6, one, one, one, false);
    }
  }
  // ---- filter factors (les.f90:203-222, 499-518, 775-799)
  // per-plane factors, formed in FP64 from the FP64 metrics
  double dyf_f, dy_f;
  {
    const double ddxs = a.dx1 * kBetaSgs, ddzs = a.dz1 * kBetaSgs, ddy = a.dy[j], ddyf = a.dyf[j];
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
  }
  const R dyf_fac = (R)dyf_f, dy_fac = (R)dy_f;
  // GYF stresses (Sij2 is handled implicitly through nu_t)
  const R f1 = (gyf_in && have_up) ? s1 : one;
  a.tau11[o] = dyf_fac * (f1 * S11);
  a.tau33[o] = dyf_fac * (f1 * S33);
  a.tau13[o] = dyf_fac * (f1 * S13);
  // GY stresses, less the parts treated implicitly in the CN solve
  const R du1 = (a.u1[o] - a.u1[o - pp]) / dyj, du3 = (a.u3[o] - a.u3[o - pp]) / dyj;
  if (gy_in) {
    a.tau12[o] = dy_fac * (temp * (S12 - h5 * du1));
    a.tau23[o] = dy_fac * (temp * (S23 - h5 * du3));
  } else {
    a.tau12[o] = dy_fac * S12;
    a.tau23[o] = dy_fac * S23;
  }
  // eddy viscosity at GY points, zero at planes 1, 2, Ny, Ny+1 (ghost_les_mpi)
  R nt = zero;
  if (j >= 3 && j <= ny - 1 && gy_in) nt = -h5 * dy_fac * temp;
  a.nu_t[o] = (double)nt;
  a.cr1[o] = (j >= 2 && j <= ny) ? nt * du1 : zero;
  a.cr3[o] = (j >= 2 && j <= ny) ? nt * du3 : zero;
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

// Types 1/4: kappa_t = nu_t and the explicit horizontal diffusion fluxes
// kappa_t d1 th, kappa_t d3 th (les.f90:1142-1203).  Type 5: the AMD scalar
// model at GYF (:1238-1275) and GY points (:1278-1330), kappa_t (:1373-1384)
// and the fluxes 0.5 delta_yf s1_th d_i th (:1334-1342, 1392-1401), with the
// filter factor applied before the forward transform.
template <typename R>
__global__ void __launch_bounds__(256) k_les_th(LesThArgs<R> a, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int j = 1 + (int)(idx / pp);
  if (j > g.ny + 1) return;
  const long long o = (long long)j * pp + idx % pp;
  const int ny = g.ny;
  const R h5 = (R)0.5, two = (R)2, zero = (R)0;
  if (a.type != 5) {
    const double kt = a.nu_t[o];
    a.kappa_t[o] = kt;
    a.f1[o] = (R)kt * a.T1[o];
    a.f3[o] = (R)kt * a.T3[o];
    return;
  }
  const R* G11 = a.G[0];
  const R* G12 = a.G[1];
  const R* G13 = a.G[2];
  const R* G21 = a.G[3];
  const R* G22 = a.G[4];
  const R* G23 = a.G[5];
  const R* G31 = a.G[6];
  const R* G32 = a.G[7];
  const R* G33 = a.G[8];
  const double ddxs = a.dx1 * kBetaSgs, ddzs = a.dz1 * kBetaSgs;
  const R dxs = (R)ddxs, dzs = (R)ddzs;
  const R dyj = (R)a.dy[j], dyfj = (R)a.dyf[j], dyfm = (R)a.dyf[j - 1];
  const double c2 = -2.0 * kCAmd * kCAmd * 3.0;
  // GYF point: du2dy(j) of the reference is G22(j-1)
  R f1 = zero, f3 = zero;
  if (j >= a.jst && j <= a.jet && j <= ny) {
    const long long u = o + pp, m = o - pp;
    const R dly = dyfj * two;
    const R dx2 = dxs * dxs, dy2 = dly * dly, dz2 = dzs * dzs;
    const R tx = a.T1[o], tz = a.T3[o], ty = h5 * (a.T2[u] + a.T2[o]);
    R v = -(dx2 * G11[o] * tx * tx + dy2 * G22[m] * ty * ty + dz2 * G33[o] * tz * tz +
            (dy2 * h5 * (G12[u] + G12[o]) + dx2 * h5 * (G21[u] + G21[o])) * tx * ty +
            (dz2 * G13[o] + dx2 * G31[o]) * tx * tz +
            (dz2 * h5 * (G23[u] + G23[o]) + dy2 * h5 * (G32[u] + G32[o])) * ty * tz);
    if (v <= zero)
      v = zero;
    else
      v = v / ((dxs * tx) * (dxs * tx) + (dly * ty) * (dly * ty) + (dzs * tz) * (dzs * tz));
    const double ddly = a.dyf[j] * 2.0;
    const R dyf_fac = (R)(c2 / (1.0 / (ddxs * ddxs) + 1.0 / (ddly * ddly) + 1.0 / (ddzs * ddzs)));
    f1 = h5 * dyf_fac * (v * tx);
    f3 = h5 * dyf_fac * (v * tz);
  }
  a.f1[o] = f1;
  a.f3[o] = f3;
  // GY point
  R kt = zero;
  if (j >= 3 && j <= ny - 1) {
    const long long m = o - pp;
    const R dly = dyj * two;
    const R dx2 = dxs * dxs, dy2 = dly * dly, dz2 = dzs * dzs;
    auto gy = [&](const R* f) { return (f[o] * dyfm + f[m] * dyfj) / (two * dyj); };
    // du2dy at GY interpolates du2dy(j) = G22(j-1) and du2dy(j-1) = G22(j-2)
    const R d22 = (G22[m] * dyfm + G22[m - pp] * dyfj) / (two * dyj);
    const R tx = gy(a.T1), tz = gy(a.T3), ty = a.T2[o];
    R v = -(dx2 * gy(G11) * tx * tx + dy2 * d22 * ty * ty + dz2 * gy(G33) * tz * tz +
            (dy2 * G12[o] + dx2 * G21[o]) * tx * ty + (dz2 * gy(G13) + dx2 * gy(G31)) * tx * tz +
            (dz2 * G23[o] + dy2 * G32[o]) * ty * tz);
    if (v <= zero)
      v = zero;
    else
      v = v / ((dxs * tx) * (dxs * tx) + (dly * ty) * (dly * ty) + (dzs * tz) * (dzs * tz));
    const double ddly = a.dy[j] * 2.0;
    const R dy_fac = (R)(c2 / (1.0 / (ddxs * ddxs) + 1.0 / (ddly * ddly) + 1.0 / (ddzs * ddzs)));
    kt = -h5 * dy_fac * v;
  }
  a.kappa_t[o] = (double)kt;
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
  const int j = (int)(idx / sp);
  if (j >= g.ny2) return;
  const long long m = idx % sp;
  const int q = (int)(m / g.nkk), kk = (int)(m % g.nkk);
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
// :1335-1496 run sequentially; b' goes to a scratch array and g' to the
// solution's own storage.
struct TriRow {
  double a, b, c;   // lane-independent part of the row (all of it for fixed rows)
  double vb;        // fixed-row right-hand side (scaled by kscale when kind == 2)
  double wp, wm;    // explicit stencil weights (assembled rows)
  double lp, lm;    // LES weights: temp2/(dy(j+1) dyf(j)), temp2/(dy(j) dyf(j)) (u type) or
                    // temp2/(dyf(j) dy(j)), temp2/(dyf(j-1) dy(j)) (v type)
  int kind;         // 1 assembled, 0 fixed, 2 fixed scaled
  int vtype;        // 1 for u2 (GY rows: nu_t averaged to GYF)
};

// T is the storage type of r and u; the sweeps run in FP64 with b' and g' in
// FP64 scratch.  incr has the meaning of k_thomas_seg's flag.
template <typename T>
__global__ void __launch_bounds__(256) k_thomas_les(const T* __restrict__ r, T* __restrict__ u,
                                                    const double* __restrict__ nt, double* __restrict__ sb,
                                                    double* __restrict__ sg, const TriRow* __restrict__ rows,
                                                    int n, long long pp, int ca, int cb, double kscale,
                                                    int incr) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= pp) return;
  auto U = [&](int j) { return (double)u[(long long)j * pp + col]; };
  auto N = [&](int j) { return nt[(long long)j * pp + col]; };
  // row coefficients of this lane
  auto row = [&](int j, double& a, double& b, double& c, double& vec, double um, double uc, double up) {
    const TriRow& rw = rows[j];
    a = rw.a;
    b = rw.b;
    c = rw.c;
    if (rw.kind == 1) {
      double nl, nu;  // eddy viscosity on the lower / upper face of the row
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
      vec = (double)r[(long long)j * pp + col] + (incr ? uc : 0.0) + rw.wp * (up - uc) - rw.wm * (uc - um) +
            (rw.lp * nu * (up - uc) - rw.lm * nl * (uc - um));
    } else {
      vec = (rw.kind == 2) ? rw.vb * kscale : rw.vb;
    }
  };
  double um = 0.0, uc = U(0), up = (n > 1) ? U(1) : 0.0;
  double a, b, c, vec;
  row(0, a, b, c, vec, um, uc, up);
  double bp = b, gp = vec, cprev = c;
  sb[col] = bp;
  sg[col] = gp;
  for (int j = 1; j < n; ++j) {
    um = uc;
    uc = up;
    up = (j + 1 < n) ? U(j + 1) : 0.0;
    row(j, a, b, c, vec, um, uc, up);
    const double mlt = -a / bp;
    bp = b + mlt * cprev;
    gp = vec + mlt * gp;
    cprev = c;
    const long long o = (long long)j * pp + col;
    sb[o] = bp;
    sg[o] = gp;
  }
  // backward: x(n-1) = g'/b', x(j) = (g'(j) - c(j) x(j+1)) / b'(j)
  double x = gp / bp;
  if (n - 1 >= ca && n - 1 <= cb) {
    const long long o = (long long)(n - 1) * pp + col;
    u[o] = (T)(x - (incr ? (double)u[o] : 0.0));
  }
  for (int j = n - 2; j >= ca; --j) {
    const TriRow& rw = rows[j];
    double c = rw.c;
    if (rw.kind == 1) {
      const double nu = rw.vtype ? 0.5 * (N(j) + N(j + 1)) : N(j + 1);
      c -= rw.lp * nu;
    }
    const long long o = (long long)j * pp + col;
    x = (sg[o] - c * x) / sb[o];
    if (j <= cb) u[o] = (T)(x - (incr ? (double)u[o] : 0.0));
  }
}

}  // namespace dgpu
