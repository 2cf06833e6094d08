// les_point.h -- the pointwise subgrid closures of les.f90 and their
// vector-Jacobian products (for the adjoint substep).
//
// Plain C++ (host and device): the velocity model of les_chan
// (les.f90:94-270 Smagorinsky, :279-565 AMD-Rozema, :570-850
// AMD-Verstappen) and the scalar AMD model of les_chan_th (:1236-1384) at
// one physical point, written on the nine velocity-gradient components
// G_ab = d_b u_a of the point's plane and of its neighbouring planes (the
// layout of the gradient fields is described in dgpu_les.h).  The VJPs are
// hand-derived: for outputs y = f(x) and a weight vector ybar they add
// (df/dx)^T ybar to xbar.  The clip max(0, .) of the AMD models has
// derivative 0 on the clipped side.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define LP_HD __host__ __device__ __forceinline__
#else
#define LP_HD inline
#endif

namespace lespt {

constexpr double kBetaSgs = 3.0;  // les.f90:60
constexpr double kCSmag = 0.17;   // les.f90:41
constexpr double kCAmd = 0.2887;  // les.f90:42

// ------------------------------------------------------------ AMD value --
// f = num / den if num > 0 else 0, with num = -(I3 - I4) and den = I1 - I2
// for strain components a1..a6 (11 22 33 12 13 23) and rotation w4..w6
// (12 13 23); Rozema weights dxq, dyq, dzq or unit weights.
template <typename R>
struct Amd {
  R a1, a2, a3, a4, a5, a6, w4, w5, w6;
};

template <typename R>
LP_HD R amd_num(const Amd<R>& z, R dxq, R dyq, R dzq, bool rozema) {
  const R a1 = z.a1, a2 = z.a2, a3 = z.a3, a4 = z.a4, a5 = z.a5, a6 = z.a6, w4 = z.w4, w5 = z.w5, w6 = z.w6;
  R v;
  if (rozema) {
    v = dxq * a1 * a1 * a1 + dyq * a2 * a2 * a2 + dzq * a3 * a3 * a3 + (2 * dxq + dyq) * a1 * a4 * a4 +
        (2 * dxq + dzq) * a1 * a5 * a5 + (2 * dyq + dxq) * a2 * a4 * a4 + (2 * dyq + dzq) * a2 * a6 * a6 +
        (2 * dzq + dxq) * a3 * a5 * a5 + (2 * dzq + dyq) * a3 * a6 * a6 + 2 * (dxq + dyq + dzq) * a4 * a5 * a6;
    v = -v - dyq * a1 * w4 * w4 - dzq * a1 * w5 * w5 - dxq * a2 * w4 * w4 - dzq * a2 * w6 * w6 -
        dxq * a3 * w5 * w5 - dyq * a3 * w6 * w6 - 2 * dzq * a4 * w5 * w6 + 2 * dyq * a5 * w4 * w6 -
        2 * dxq * a6 * w4 * w5;
  } else {
    v = a1 * a1 * a1 + a2 * a2 * a2 + a3 * a3 * a3 + 3 * a1 * a4 * a4 + 3 * a1 * a5 * a5 + 3 * a2 * a4 * a4 +
        3 * a2 * a6 * a6 + 3 * a3 * a5 * a5 + 3 * a3 * a6 * a6 + 6 * a4 * a5 * a6;
    v = -v - a1 * w4 * w4 - a1 * w5 * w5 - a2 * w4 * w4 - a2 * w6 * w6 - a3 * w5 * w5 - a3 * w6 * w6 -
        2 * a4 * w5 * w6 + 2 * a5 * w4 * w6 - 2 * a6 * w4 * w5;
  }
  return v;
}

template <typename R>
LP_HD R amd_den(const Amd<R>& z) {
  return z.a1 * z.a1 + z.a2 * z.a2 + z.a3 * z.a3 + 2 * z.a4 * z.a4 + 2 * z.w4 * z.w4 + 2 * z.a5 * z.a5 +
         2 * z.w5 * z.w5 + 2 * z.a6 * z.a6 + 2 * z.w6 * z.w6;
}

template <typename R>
LP_HD R amd_value(const Amd<R>& z, R dxq, R dyq, R dzq, bool rozema) {
  const R v = amd_num(z, dxq, dyq, dzq, rozema);
  if (v <= (R)0) return (R)0;
  return v / amd_den(z);
}

// zbar += (d amd_value / dz)^T fbar
LP_HD void amd_vjp(const Amd<double>& z, double dxq, double dyq, double dzq, bool rozema, double fbar,
                          Amd<double>& zb) {
  const double num = amd_num(z, dxq, dyq, dzq, rozema);
  if (num <= 0.0 || fbar == 0.0) return;
  const double den = amd_den(z);
  const double f = num / den;
  const double a1 = z.a1, a2 = z.a2, a3 = z.a3, a4 = z.a4, a5 = z.a5, a6 = z.a6, w4 = z.w4, w5 = z.w5,
               w6 = z.w6;
  double n1, n2, n3, n4, n5, n6, m4, m5, m6;  // d num / d (a1..a6, w4..w6)
  if (rozema) {
    const double cx = 2 * dxq, cy = 2 * dyq, cz = 2 * dzq, c3 = 2 * (dxq + dyq + dzq);
    n1 = -(3 * dxq * a1 * a1 + (cx + dyq) * a4 * a4 + (cx + dzq) * a5 * a5) - dyq * w4 * w4 - dzq * w5 * w5;
    n2 = -(3 * dyq * a2 * a2 + (cy + dxq) * a4 * a4 + (cy + dzq) * a6 * a6) - dxq * w4 * w4 - dzq * w6 * w6;
    n3 = -(3 * dzq * a3 * a3 + (cz + dxq) * a5 * a5 + (cz + dyq) * a6 * a6) - dxq * w5 * w5 - dyq * w6 * w6;
    n4 = -(2 * (cx + dyq) * a1 * a4 + 2 * (cy + dxq) * a2 * a4 + c3 * a5 * a6) - 2 * dzq * w5 * w6;
    n5 = -(2 * (cx + dzq) * a1 * a5 + 2 * (cz + dxq) * a3 * a5 + c3 * a4 * a6) + 2 * dyq * w4 * w6;
    n6 = -(2 * (cy + dzq) * a2 * a6 + 2 * (cz + dyq) * a3 * a6 + c3 * a4 * a5) - 2 * dxq * w4 * w5;
    m4 = -2 * dyq * a1 * w4 - 2 * dxq * a2 * w4 + 2 * dyq * a5 * w6 - 2 * dxq * a6 * w5;
    m5 = -2 * dzq * a1 * w5 - 2 * dxq * a3 * w5 - 2 * dzq * a4 * w6 - 2 * dxq * a6 * w4;
    m6 = -2 * dzq * a2 * w6 - 2 * dyq * a3 * w6 - 2 * dzq * a4 * w5 + 2 * dyq * a5 * w4;
  } else {
    n1 = -(3 * a1 * a1 + 3 * a4 * a4 + 3 * a5 * a5) - w4 * w4 - w5 * w5;
    n2 = -(3 * a2 * a2 + 3 * a4 * a4 + 3 * a6 * a6) - w4 * w4 - w6 * w6;
    n3 = -(3 * a3 * a3 + 3 * a5 * a5 + 3 * a6 * a6) - w5 * w5 - w6 * w6;
    n4 = -(6 * a1 * a4 + 6 * a2 * a4 + 6 * a5 * a6) - 2 * w5 * w6;
    n5 = -(6 * a1 * a5 + 6 * a3 * a5 + 6 * a4 * a6) + 2 * w4 * w6;
    n6 = -(6 * a2 * a6 + 6 * a3 * a6 + 6 * a4 * a5) - 2 * w4 * w5;
    m4 = -2 * a1 * w4 - 2 * a2 * w4 + 2 * a5 * w6 - 2 * a6 * w5;
    m5 = -2 * a1 * w5 - 2 * a3 * w5 - 2 * a4 * w6 - 2 * a6 * w4;
    m6 = -2 * a2 * w6 - 2 * a3 * w6 - 2 * a4 * w5 + 2 * a5 * w4;
  }
  // d f = (d num - f d den) / den,  d den = 2 a1, 2 a2, 2 a3, 4 a4, 4 a5, 4 a6, 4 w4, 4 w5, 4 w6
  const double s = fbar / den;
  zb.a1 += s * (n1 - f * 2 * a1);
  zb.a2 += s * (n2 - f * 2 * a2);
  zb.a3 += s * (n3 - f * 2 * a3);
  zb.a4 += s * (n4 - f * 4 * a4);
  zb.a5 += s * (n5 - f * 4 * a5);
  zb.a6 += s * (n6 - f * 4 * a6);
  zb.w4 += s * (m4 - f * 4 * w4);
  zb.w5 += s * (m5 - f * 4 * w5);
  zb.w6 += s * (m6 - f * 4 * w6);
}

// ----------------------------------------------------- velocity model --
// Inputs at one point of plane j: the nine gradients c[] at the point, four
// at plane j+1 (G12 G21 G23 G32), five at plane j-1 (G11 G22 G33 G13 G31),
// and u1, u3 at the point and at plane j-1.
template <typename R>
struct VelIn {
  R c[9];   // G11 G12 G13 G21 G22 G23 G31 G32 G33
  R up[4];  // plane j+1: G12 G21 G23 G32
  R dn[5];  // plane j-1: G11 G22 G33 G13 G31
  R u1o, u1m, u3o, u3m;
};

template <typename R>
struct VelOut {
  R t11, t33, t13, t12, t23, nt, cr1, cr3;
};

// Per-plane constants and masks of the velocity model (k_les_model).
template <typename R>
struct VelPlane {
  int type;
  bool gyf_in, gy_in, nt_in, cr_in;
  R dyj, dyfj, dyfm, dysu;  // dy(j), dyf(j), dyf(j-1), 2 dy(j+1)
  R dxs, dzs;               // beta_sgs Lx/Nx, beta_sgs Lz/Nz
  R dyf_fac, dy_fac;        // filter factors of the GYF and GY stresses
};

enum { G11 = 0, G12, G13, G21, G22, G23, G31, G32, G33 };
enum { U12 = 0, U21, U23, U32 };
enum { D11 = 0, D22, D33, D13, D31 };

template <typename R>
LP_HD R gyi(R o, R m, const VelPlane<R>& P) {
  return (o * P.dyfm + m * P.dyfj) / ((R)2 * P.dyj);
}

// The model's scalar at the GYF point: |S| (type 1) or the AMD value.
template <typename R>
LP_HD R vel_gyf_scalar(const VelPlane<R>& P, const VelIn<R>& x, Amd<R>* zout = nullptr) {
  const R h5 = (R)0.5, two = (R)2, four = (R)4, one = (R)1;
  const R* c = x.c;
  const R* u = x.up;
  const R S11 = c[G11], S22 = c[G22], S33 = c[G33];
  const R S13 = h5 * (c[G13] + c[G31]);
  const R S12 = h5 * (c[G12] + c[G21]);
  const R S23 = h5 * (c[G23] + c[G32]);
  if (P.type == 1) {
    const R s12 = h5 * (h5 * (u[U12] + u[U21]) + S12);
    const R s23 = h5 * (h5 * (u[U23] + u[U32]) + S23);
    return sqrt(two * S11 * S11 + four * s12 * s12 + four * S13 * S13 + two * S22 * S22 + four * s23 * s23 +
                two * S33 * S33);
  }
  Amd<R> z;
  R dxq = one, dyq = one, dzq = one;
  if (P.type == 4) {
    dxq = P.dxs * P.dxs;
    dzq = P.dzs * P.dzs;
    dyq = (P.dyj * two) * (P.dyj * two);
    z = {S11, S22, S33, h5 * (h5 * (u[U12] + u[U21]) + S12), S13, h5 * (h5 * (u[U23] + u[U32]) + S23),
         h5 * (h5 * (u[U12] - u[U21]) + h5 * (c[G12] - c[G21])), h5 * (c[G13] - c[G31]),
         h5 * (h5 * (u[U23] - u[U32]) + h5 * (c[G23] - c[G32]))};
  } else {
    const R dxs = P.dxs, dzs = P.dzs, dysu = P.dysu, dysj = P.dyj * two;
    const R si4u = h5 * (u[U12] * (dysu / dxs) + u[U21] * (dxs / dysu));
    const R si4j = h5 * (c[G12] * (dysj / dxs) + c[G21] * (dxs / dysj));
    const R si6u = h5 * (u[U23] * (dzs / dysu) + u[U32] * (dysu / dzs));
    const R si6j = h5 * (c[G23] * (dzs / dysj) + c[G32] * (dysj / dzs));
    const R o4u = h5 * (u[U12] * (dysu / dxs) - u[U21] * (dxs / dysu));
    const R o4j = h5 * (c[G12] * (dysj / dxs) - c[G21] * (dxs / dysj));
    const R o6u = h5 * (u[U23] * (dzs / dysu) - u[U32] * (dysu / dzs));
    const R o6j = h5 * (c[G23] * (dzs / dysj) - c[G32] * (dysj / dzs));
    const R si5 = h5 * (c[G13] * (dzs / dxs) + c[G31] * (dxs / dzs));
    const R o5 = h5 * (c[G13] * (dzs / dxs) - c[G31] * (dxs / dzs));
    z = {S11, S22, S33, h5 * (si4u + si4j), si5, h5 * (si6u + si6j), h5 * (o4u + o4j), o5, h5 * (o6u + o6j)};
  }
  if (zout) *zout = z;
  return amd_value(z, dxq, dyq, dzq, P.type == 4);
}

// The model's scalar at the GY point.
template <typename R>
LP_HD R vel_gy_scalar(const VelPlane<R>& P, const VelIn<R>& x, Amd<R>* zout = nullptr) {
  const R h5 = (R)0.5, two = (R)2, four = (R)4, one = (R)1;
  const R* c = x.c;
  const R* d = x.dn;
  const R S12 = h5 * (c[G12] + c[G21]);
  const R S23 = h5 * (c[G23] + c[G32]);
  const R i11 = gyi(c[G11], d[D11], P), i22 = gyi(c[G22], d[D22], P), i33 = gyi(c[G33], d[D33], P);
  if (P.type == 1) {
    const R i13 = h5 * (gyi(c[G13], d[D13], P) + gyi(c[G31], d[D31], P));
    return sqrt(two * i11 * i11 + four * S12 * S12 + four * i13 * i13 + two * i22 * i22 + four * S23 * S23 +
                two * i33 * i33);
  }
  Amd<R> z;
  R dxq = one, dyq = one, dzq = one;
  if (P.type == 4) {
    dxq = P.dxs * P.dxs;
    dzq = P.dzs * P.dzs;
    dyq = (P.dyj * two) * (P.dyj * two);
    const R i13 = h5 * (gyi(c[G13], d[D13], P) + gyi(c[G31], d[D31], P));
    const R w5 = (h5 * (c[G13] - c[G31]) * P.dyfm + h5 * (d[D13] - d[D31]) * P.dyfj) / (two * P.dyj);
    z = {i11, i22, i33, S12, i13, S23, h5 * (c[G12] - c[G21]), w5, h5 * (c[G23] - c[G32])};
  } else {
    const R dxs = P.dxs, dzs = P.dzs, dysj = P.dyj * two;
    const R si4 = h5 * (c[G12] * (dysj / dxs) + c[G21] * (dxs / dysj));
    const R si6 = h5 * (c[G23] * (dzs / dysj) + c[G32] * (dysj / dzs));
    const R o4 = h5 * (c[G12] * (dysj / dxs) - c[G21] * (dxs / dysj));
    const R o6 = h5 * (c[G23] * (dzs / dysj) - c[G32] * (dysj / dzs));
    const R si5o = h5 * (c[G13] * (dzs / dxs) + c[G31] * (dxs / dzs));
    const R si5m = h5 * (d[D13] * (dzs / dxs) + d[D31] * (dxs / dzs));
    const R o5o = h5 * (c[G13] * (dzs / dxs) - c[G31] * (dxs / dzs));
    const R o5m = h5 * (d[D13] * (dzs / dxs) - d[D31] * (dxs / dzs));
    const R si5 = (si5o * P.dyfm + si5m * P.dyfj) / (two * P.dyj);
    const R o5 = (o5o * P.dyfm + o5m * P.dyfj) / (two * P.dyj);
    z = {i11, i22, i33, si4, si5, si6, o4, o5, o6};
  }
  if (zout) *zout = z;
  return amd_value(z, dxq, dyq, dzq, P.type == 4);
}

// All outputs of the velocity model (k_les_model) at one point (les.f90:201-222, 498-518, 773-799,
// 890-919): filter-scaled stresses, nu_t, and the cross-term fluxes.
template <typename R>
LP_HD void vel_point(const VelPlane<R>& P, const VelIn<R>& x, VelOut<R>& y) {
  const R h5 = (R)0.5, one = (R)1, zero = (R)0;
  const R* c = x.c;
  const R S11 = c[G11], S33 = c[G33];
  const R S13 = h5 * (c[G13] + c[G31]);
  const R S12 = h5 * (c[G12] + c[G21]);
  const R S23 = h5 * (c[G23] + c[G32]);
  const R s1 = P.gyf_in ? vel_gyf_scalar(P, x) : zero;
  const R temp = P.gy_in ? vel_gy_scalar(P, x) : zero;
  const R f1 = P.gyf_in ? s1 : one;
  y.t11 = P.dyf_fac * (f1 * S11);
  y.t33 = P.dyf_fac * (f1 * S33);
  y.t13 = P.dyf_fac * (f1 * S13);
  const R du1 = (x.u1o - x.u1m) / P.dyj, du3 = (x.u3o - x.u3m) / P.dyj;
  if (P.gy_in) {
    y.t12 = P.dy_fac * (temp * (S12 - h5 * du1));
    y.t23 = P.dy_fac * (temp * (S23 - h5 * du3));
  } else {
    y.t12 = P.dy_fac * S12;
    y.t23 = P.dy_fac * S23;
  }
  const R nt = P.nt_in ? -h5 * P.dy_fac * temp : zero;
  y.nt = nt;
  y.cr1 = P.cr_in ? nt * du1 : zero;
  y.cr3 = P.cr_in ? nt * du3 : zero;
}

// d(sqrt(Q))/d(inputs) helpers for Smagorinsky
LP_HD double safe_half_inv(double s) { return s > 0.0 ? 0.5 / s : 0.0; }

// xbar += (d vel_gyf_scalar / dx)^T sbar
LP_HD void vel_gyf_scalar_vjp(const VelPlane<double>& P, const VelIn<double>& x, double sbar,
                                     VelIn<double>& xb) {
  if (sbar == 0.0) return;
  const double h5 = 0.5;
  const double* c = x.c;
  const double* u = x.up;
  double* cb = xb.c;
  double* ub = xb.up;
  if (P.type == 1) {
    const double S11 = c[G11], S22 = c[G22], S33 = c[G33], S13 = h5 * (c[G13] + c[G31]);
    const double S12 = h5 * (c[G12] + c[G21]), S23 = h5 * (c[G23] + c[G32]);
    const double s12 = h5 * (h5 * (u[U12] + u[U21]) + S12);
    const double s23 = h5 * (h5 * (u[U23] + u[U32]) + S23);
    const double s = sqrt(2 * S11 * S11 + 4 * s12 * s12 + 4 * S13 * S13 + 2 * S22 * S22 + 4 * s23 * s23 +
                          2 * S33 * S33);
    const double q = sbar * safe_half_inv(s);  // dQ weight
    cb[G11] += q * 4 * S11;
    cb[G22] += q * 4 * S22;
    cb[G33] += q * 4 * S33;
    const double b13 = q * 8 * S13;
    cb[G13] += h5 * b13;
    cb[G31] += h5 * b13;
    const double b12 = q * 8 * s12, b23 = q * 8 * s23;  // d/ds12, d/ds23
    // s12 = 1/4 (u12 + u21) + 1/4 (c12 + c21)
    ub[U12] += 0.25 * b12;
    ub[U21] += 0.25 * b12;
    cb[G12] += 0.25 * b12;
    cb[G21] += 0.25 * b12;
    ub[U23] += 0.25 * b23;
    ub[U32] += 0.25 * b23;
    cb[G23] += 0.25 * b23;
    cb[G32] += 0.25 * b23;
    return;
  }
  Amd<double> z;
  vel_gyf_scalar(P, x, &z);
  double dxq = 1, dyq = 1, dzq = 1;
  if (P.type == 4) {
    dxq = P.dxs * P.dxs;
    dzq = P.dzs * P.dzs;
    dyq = (P.dyj * 2) * (P.dyj * 2);
  }
  Amd<double> zb = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  amd_vjp(z, dxq, dyq, dzq, P.type == 4, sbar, zb);
  cb[G11] += zb.a1;
  cb[G22] += zb.a2;
  cb[G33] += zb.a3;
  if (P.type == 4) {
    // a4 = 1/4 (u12 + u21 + c12 + c21), a6 likewise, a5 = 1/2 (c13 + c31),
    // w4 = 1/4 (u12 - u21 + c12 - c21), w5 = 1/2 (c13 - c31), w6 = 1/4 (u23 - u32 + c23 - c32)
    ub[U12] += 0.25 * (zb.a4 + zb.w4);
    ub[U21] += 0.25 * (zb.a4 - zb.w4);
    cb[G12] += 0.25 * (zb.a4 + zb.w4);
    cb[G21] += 0.25 * (zb.a4 - zb.w4);
    ub[U23] += 0.25 * (zb.a6 + zb.w6);
    ub[U32] += 0.25 * (zb.a6 - zb.w6);
    cb[G23] += 0.25 * (zb.a6 + zb.w6);
    cb[G32] += 0.25 * (zb.a6 - zb.w6);
    cb[G13] += 0.5 * (zb.a5 + zb.w5);
    cb[G31] += 0.5 * (zb.a5 - zb.w5);
  } else {
    const double dxs = P.dxs, dzs = P.dzs, dysu = P.dysu, dysj = P.dyj * 2;
    // a4 = 1/2 (si4u + si4j), w4 = 1/2 (o4u + o4j); si4u = 1/2 (u12 A + u21 B), o4u = 1/2 (u12 A - u21 B)
    const double A4u = dysu / dxs, B4u = dxs / dysu, A4j = dysj / dxs, B4j = dxs / dysj;
    const double A6u = dzs / dysu, B6u = dysu / dzs, A6j = dzs / dysj, B6j = dysj / dzs;
    const double p4 = 0.25 * zb.a4, q4 = 0.25 * zb.w4, p6 = 0.25 * zb.a6, q6 = 0.25 * zb.w6;
    ub[U12] += (p4 + q4) * A4u;
    ub[U21] += (p4 - q4) * B4u;
    cb[G12] += (p4 + q4) * A4j;
    cb[G21] += (p4 - q4) * B4j;
    ub[U23] += (p6 + q6) * A6u;
    ub[U32] += (p6 - q6) * B6u;
    cb[G23] += (p6 + q6) * A6j;
    cb[G32] += (p6 - q6) * B6j;
    // a5 = si5 = 1/2 (c13 A5 + c31 B5), w5 = o5 = 1/2 (c13 A5 - c31 B5)
    const double A5 = dzs / dxs, B5 = dxs / dzs;
    cb[G13] += 0.5 * (zb.a5 + zb.w5) * A5;
    cb[G31] += 0.5 * (zb.a5 - zb.w5) * B5;
  }
}

// xbar += (d vel_gy_scalar / dx)^T tbar
LP_HD void vel_gy_scalar_vjp(const VelPlane<double>& P, const VelIn<double>& x, double tbar,
                                    VelIn<double>& xb) {
  if (tbar == 0.0) return;
  const double h5 = 0.5;
  const double* c = x.c;
  const double* d = x.dn;
  double* cb = xb.c;
  double* db = xb.dn;
  const double wo = P.dyfm / (2 * P.dyj), wm = P.dyfj / (2 * P.dyj);  // gyi weights
  if (P.type == 1) {
    const double S12 = h5 * (c[G12] + c[G21]), S23 = h5 * (c[G23] + c[G32]);
    const double i11 = gyi(c[G11], d[D11], P), i22 = gyi(c[G22], d[D22], P), i33 = gyi(c[G33], d[D33], P);
    const double i13 = h5 * (gyi(c[G13], d[D13], P) + gyi(c[G31], d[D31], P));
    const double t = sqrt(2 * i11 * i11 + 4 * S12 * S12 + 4 * i13 * i13 + 2 * i22 * i22 + 4 * S23 * S23 +
                          2 * i33 * i33);
    const double q = tbar * safe_half_inv(t);
    const double b11 = q * 4 * i11, b22 = q * 4 * i22, b33 = q * 4 * i33, b13 = q * 8 * i13;
    cb[G11] += b11 * wo;
    db[D11] += b11 * wm;
    cb[G22] += b22 * wo;
    db[D22] += b22 * wm;
    cb[G33] += b33 * wo;
    db[D33] += b33 * wm;
    cb[G13] += h5 * b13 * wo;
    db[D13] += h5 * b13 * wm;
    cb[G31] += h5 * b13 * wo;
    db[D31] += h5 * b13 * wm;
    const double b12 = q * 8 * S12, b23 = q * 8 * S23;
    cb[G12] += h5 * b12;
    cb[G21] += h5 * b12;
    cb[G23] += h5 * b23;
    cb[G32] += h5 * b23;
    return;
  }
  Amd<double> z;
  vel_gy_scalar(P, x, &z);
  double dxq = 1, dyq = 1, dzq = 1;
  if (P.type == 4) {
    dxq = P.dxs * P.dxs;
    dzq = P.dzs * P.dzs;
    dyq = (P.dyj * 2) * (P.dyj * 2);
  }
  Amd<double> zb = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  amd_vjp(z, dxq, dyq, dzq, P.type == 4, tbar, zb);
  // a1..a3 = i11, i22, i33
  cb[G11] += zb.a1 * wo;
  db[D11] += zb.a1 * wm;
  cb[G22] += zb.a2 * wo;
  db[D22] += zb.a2 * wm;
  cb[G33] += zb.a3 * wo;
  db[D33] += zb.a3 * wm;
  if (P.type == 4) {
    // a4 = S12, a6 = S23, a5 = i13, w4 = 1/2 (c12 - c21), w6 = 1/2 (c23 - c32),
    // w5 = 1/2 (c13 - c31) wo + 1/2 (d13 - d31) wm
    cb[G12] += h5 * (zb.a4 + zb.w4);
    cb[G21] += h5 * (zb.a4 - zb.w4);
    cb[G23] += h5 * (zb.a6 + zb.w6);
    cb[G32] += h5 * (zb.a6 - zb.w6);
    cb[G13] += h5 * (zb.a5 + zb.w5) * wo;
    cb[G31] += h5 * (zb.a5 - zb.w5) * wo;
    db[D13] += h5 * (zb.a5 + zb.w5) * wm;
    db[D31] += h5 * (zb.a5 - zb.w5) * wm;
  } else {
    const double dxs = P.dxs, dzs = P.dzs, dysj = P.dyj * 2;
    const double A4 = dysj / dxs, B4 = dxs / dysj, A6 = dzs / dysj, B6 = dysj / dzs;
    const double A5 = dzs / dxs, B5 = dxs / dzs;
    cb[G12] += h5 * (zb.a4 + zb.w4) * A4;
    cb[G21] += h5 * (zb.a4 - zb.w4) * B4;
    cb[G23] += h5 * (zb.a6 + zb.w6) * A6;
    cb[G32] += h5 * (zb.a6 - zb.w6) * B6;
    // a5 = si5o wo + si5m wm, w5 = o5o wo + o5m wm
    cb[G13] += h5 * (zb.a5 + zb.w5) * A5 * wo;
    cb[G31] += h5 * (zb.a5 - zb.w5) * B5 * wo;
    db[D13] += h5 * (zb.a5 + zb.w5) * A5 * wm;
    db[D31] += h5 * (zb.a5 - zb.w5) * B5 * wm;
  }
}

// xbar += (d vel_point / dx)^T ybar
LP_HD void vel_point_vjp(const VelPlane<double>& P, const VelIn<double>& x, const VelOut<double>& yb,
                                VelIn<double>& xb) {
  const double h5 = 0.5;
  const double* c = x.c;
  const double S11 = c[G11], S33 = c[G33], S13 = h5 * (c[G13] + c[G31]);
  const double S12 = h5 * (c[G12] + c[G21]), S23 = h5 * (c[G23] + c[G32]);
  const double s1 = P.gyf_in ? vel_gyf_scalar(P, x) : 0.0;
  const double temp = P.gy_in ? vel_gy_scalar(P, x) : 0.0;
  const double f1 = P.gyf_in ? s1 : 1.0;
  const double du1 = (x.u1o - x.u1m) / P.dyj, du3 = (x.u3o - x.u3m) / P.dyj;
  const double nt = P.nt_in ? -h5 * P.dy_fac * temp : 0.0;
  double ntb = yb.nt, du1b = 0, du3b = 0, tempb = 0, f1b = 0;
  double S11b = 0, S33b = 0, S13b = 0, S12b = 0, S23b = 0;
  if (P.cr_in) {
    ntb += yb.cr1 * du1 + yb.cr3 * du3;
    du1b += yb.cr1 * nt;
    du3b += yb.cr3 * nt;
  }
  if (P.nt_in) tempb += -h5 * P.dy_fac * ntb;
  if (P.gy_in) {
    tempb += yb.t12 * P.dy_fac * (S12 - h5 * du1) + yb.t23 * P.dy_fac * (S23 - h5 * du3);
    S12b += yb.t12 * P.dy_fac * temp;
    S23b += yb.t23 * P.dy_fac * temp;
    du1b += -h5 * yb.t12 * P.dy_fac * temp;
    du3b += -h5 * yb.t23 * P.dy_fac * temp;
  } else {
    S12b += yb.t12 * P.dy_fac;
    S23b += yb.t23 * P.dy_fac;
  }
  f1b += P.dyf_fac * (yb.t11 * S11 + yb.t33 * S33 + yb.t13 * S13);
  S11b += P.dyf_fac * f1 * yb.t11;
  S33b += P.dyf_fac * f1 * yb.t33;
  S13b += P.dyf_fac * f1 * yb.t13;
  double* cb = xb.c;
  cb[G11] += S11b;
  cb[G33] += S33b;
  cb[G13] += h5 * S13b;
  cb[G31] += h5 * S13b;
  cb[G12] += h5 * S12b;
  cb[G21] += h5 * S12b;
  cb[G23] += h5 * S23b;
  cb[G32] += h5 * S23b;
  xb.u1o += du1b / P.dyj;
  xb.u1m -= du1b / P.dyj;
  xb.u3o += du3b / P.dyj;
  xb.u3m -= du3b / P.dyj;
  if (P.gyf_in) vel_gyf_scalar_vjp(P, x, f1b, xb);
  if (P.gy_in) vel_gy_scalar_vjp(P, x, tempb, xb);
}

// ------------------------------------------------------- scalar model --
// AMD scalar model (type 5) at one point of plane j: inputs at planes
// j-2..j+1, outputs the two horizontal fluxes (GYF) and kappa_t (GY).
template <typename R>
struct ThIn {
  R c[9];   // gradients at plane j (G22 unused)
  R t[3];   // T1 T2 T3 at plane j
  R up[5];  // plane j+1: G12 G21 G23 G32 T2
  R dn[7];  // plane j-1: G11 G22 G33 G13 G31 T1 T3
  R dd;     // plane j-2: G22
};
template <typename R>
struct ThOut {
  R f1, f3, kt;
};
template <typename R>
struct ThPlane {
  bool gyf_in, gy_in;
  R dxs, dzs, dyj, dyfj, dyfm;
  R dyf_fac, dy_fac;
};
enum { TU12 = 0, TU21, TU23, TU32, TUT2 };
enum { TD11 = 0, TD22, TD33, TD13, TD31, TDT1, TDT3 };

template <typename R>
LP_HD void th_point(const ThPlane<R>& P, const ThIn<R>& x, ThOut<R>& y) {
  const R h5 = (R)0.5, two = (R)2, zero = (R)0;
  const R* c = x.c;
  const R* u = x.up;
  const R* d = x.dn;
  y.f1 = zero;
  y.f3 = zero;
  y.kt = zero;
  if (P.gyf_in) {
    const R dly = P.dyfj * two;
    const R dx2 = P.dxs * P.dxs, dy2 = dly * dly, dz2 = P.dzs * P.dzs;
    const R tx = x.t[0], tz = x.t[2], ty = h5 * (u[TUT2] + x.t[1]);
    R v = -(dx2 * c[G11] * tx * tx + dy2 * d[TD22] * ty * ty + dz2 * c[G33] * tz * tz +
            (dy2 * h5 * (u[TU12] + c[G12]) + dx2 * h5 * (u[TU21] + c[G21])) * tx * ty +
            (dz2 * c[G13] + dx2 * c[G31]) * tx * tz +
            (dz2 * h5 * (u[TU23] + c[G23]) + dy2 * h5 * (u[TU32] + c[G32])) * ty * tz);
    if (v <= zero)
      v = zero;
    else
      v = v / ((P.dxs * tx) * (P.dxs * tx) + (dly * ty) * (dly * ty) + (P.dzs * tz) * (P.dzs * tz));
    y.f1 = h5 * P.dyf_fac * (v * tx);
    y.f3 = h5 * P.dyf_fac * (v * tz);
  }
  if (P.gy_in) {
    const R dly = P.dyj * two;
    const R dx2 = P.dxs * P.dxs, dy2 = dly * dly, dz2 = P.dzs * P.dzs;
    auto gy = [&](R o, R m) { return (o * P.dyfm + m * P.dyfj) / (two * P.dyj); };
    const R d22 = (d[TD22] * P.dyfm + x.dd * P.dyfj) / (two * P.dyj);
    const R tx = gy(x.t[0], d[TDT1]), tz = gy(x.t[2], d[TDT3]), ty = x.t[1];
    R v = -(dx2 * gy(c[G11], d[TD11]) * tx * tx + dy2 * d22 * ty * ty + dz2 * gy(c[G33], d[TD33]) * tz * tz +
            (dy2 * c[G12] + dx2 * c[G21]) * tx * ty + (dz2 * gy(c[G13], d[TD13]) + dx2 * gy(c[G31], d[TD31])) * tx * tz +
            (dz2 * c[G23] + dy2 * c[G32]) * ty * tz);
    if (v <= zero)
      v = zero;
    else
      v = v / ((P.dxs * tx) * (P.dxs * tx) + (dly * ty) * (dly * ty) + (P.dzs * tz) * (P.dzs * tz));
    y.kt = -h5 * P.dy_fac * v;
  }
}

// One AMD scalar value -(sum_ab g_ab t_a t_b * wa)/(sum (w t)^2) clipped,
// written on the 9 coefficients and the 3 scalar gradients:
//   num = -(k11 tx^2 + k22 ty^2 + k33 tz^2 + k12 tx ty + k13 tx tz + k23 ty tz)
//   den = ex tx^2 + ey ty^2 + ez tz^2
struct ThAmd {
  double k11, k22, k33, k12, k13, k23, tx, ty, tz;
};
LP_HD double th_amd_value(const ThAmd& s, double ex, double ey, double ez) {
  const double num = -(s.k11 * s.tx * s.tx + s.k22 * s.ty * s.ty + s.k33 * s.tz * s.tz + s.k12 * s.tx * s.ty +
                       s.k13 * s.tx * s.tz + s.k23 * s.ty * s.tz);
  if (num <= 0.0) return 0.0;
  return num / (ex * s.tx * s.tx + ey * s.ty * s.ty + ez * s.tz * s.tz);
}
LP_HD void th_amd_vjp(const ThAmd& s, double ex, double ey, double ez, double fb, ThAmd& sb) {
  const double num = -(s.k11 * s.tx * s.tx + s.k22 * s.ty * s.ty + s.k33 * s.tz * s.tz + s.k12 * s.tx * s.ty +
                       s.k13 * s.tx * s.tz + s.k23 * s.ty * s.tz);
  if (num <= 0.0 || fb == 0.0) return;
  const double den = ex * s.tx * s.tx + ey * s.ty * s.ty + ez * s.tz * s.tz;
  const double f = num / den, w = fb / den;
  sb.k11 += -w * s.tx * s.tx;
  sb.k22 += -w * s.ty * s.ty;
  sb.k33 += -w * s.tz * s.tz;
  sb.k12 += -w * s.tx * s.ty;
  sb.k13 += -w * s.tx * s.tz;
  sb.k23 += -w * s.ty * s.tz;
  sb.tx += w * (-(2 * s.k11 * s.tx + s.k12 * s.ty + s.k13 * s.tz) - f * 2 * ex * s.tx);
  sb.ty += w * (-(2 * s.k22 * s.ty + s.k12 * s.tx + s.k23 * s.tz) - f * 2 * ey * s.ty);
  sb.tz += w * (-(2 * s.k33 * s.tz + s.k13 * s.tx + s.k23 * s.ty) - f * 2 * ez * s.tz);
}

LP_HD void th_point_vjp(const ThPlane<double>& P, const ThIn<double>& x, const ThOut<double>& yb,
                               ThIn<double>& xb) {
  const double h5 = 0.5;
  const double* c = x.c;
  const double* u = x.up;
  const double* d = x.dn;
  if (P.gyf_in) {
    const double dly = P.dyfj * 2;
    const double dx2 = P.dxs * P.dxs, dy2 = dly * dly, dz2 = P.dzs * P.dzs;
    ThAmd s;
    s.tx = x.t[0];
    s.tz = x.t[2];
    s.ty = h5 * (u[TUT2] + x.t[1]);
    s.k11 = dx2 * c[G11];
    s.k22 = dy2 * d[TD22];
    s.k33 = dz2 * c[G33];
    s.k12 = dy2 * h5 * (u[TU12] + c[G12]) + dx2 * h5 * (u[TU21] + c[G21]);
    s.k13 = dz2 * c[G13] + dx2 * c[G31];
    s.k23 = dz2 * h5 * (u[TU23] + c[G23]) + dy2 * h5 * (u[TU32] + c[G32]);
    const double ex = P.dxs * P.dxs, ey = dly * dly, ez = P.dzs * P.dzs;
    const double v = th_amd_value(s, ex, ey, ez);
    // f1 = 1/2 fac v tx, f3 = 1/2 fac v tz
    const double vb = h5 * P.dyf_fac * (yb.f1 * s.tx + yb.f3 * s.tz);
    ThAmd sb = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    sb.tx += h5 * P.dyf_fac * v * yb.f1;
    sb.tz += h5 * P.dyf_fac * v * yb.f3;
    th_amd_vjp(s, ex, ey, ez, vb, sb);
    xb.t[0] += sb.tx;
    xb.t[2] += sb.tz;
    xb.up[TUT2] += h5 * sb.ty;
    xb.t[1] += h5 * sb.ty;
    xb.c[G11] += dx2 * sb.k11;
    xb.dn[TD22] += dy2 * sb.k22;
    xb.c[G33] += dz2 * sb.k33;
    xb.up[TU12] += dy2 * h5 * sb.k12;
    xb.c[G12] += dy2 * h5 * sb.k12;
    xb.up[TU21] += dx2 * h5 * sb.k12;
    xb.c[G21] += dx2 * h5 * sb.k12;
    xb.c[G13] += dz2 * sb.k13;
    xb.c[G31] += dx2 * sb.k13;
    xb.up[TU23] += dz2 * h5 * sb.k23;
    xb.c[G23] += dz2 * h5 * sb.k23;
    xb.up[TU32] += dy2 * h5 * sb.k23;
    xb.c[G32] += dy2 * h5 * sb.k23;
  }
  if (P.gy_in && yb.kt != 0.0) {
    const double dly = P.dyj * 2;
    const double dx2 = P.dxs * P.dxs, dy2 = dly * dly, dz2 = P.dzs * P.dzs;
    const double wo = P.dyfm / (2 * P.dyj), wm = P.dyfj / (2 * P.dyj);
    auto gy = [&](double o, double m) { return (o * P.dyfm + m * P.dyfj) / (2 * P.dyj); };
    ThAmd s;
    s.tx = gy(x.t[0], d[TDT1]);
    s.tz = gy(x.t[2], d[TDT3]);
    s.ty = x.t[1];
    const double d22 = (d[TD22] * P.dyfm + x.dd * P.dyfj) / (2 * P.dyj);
    s.k11 = dx2 * gy(c[G11], d[TD11]);
    s.k22 = dy2 * d22;
    s.k33 = dz2 * gy(c[G33], d[TD33]);
    s.k12 = dy2 * c[G12] + dx2 * c[G21];
    s.k13 = dz2 * gy(c[G13], d[TD13]) + dx2 * gy(c[G31], d[TD31]);
    s.k23 = dz2 * c[G23] + dy2 * c[G32];
    const double ex = P.dxs * P.dxs, ey = dly * dly, ez = P.dzs * P.dzs;
    ThAmd sb = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    th_amd_vjp(s, ex, ey, ez, -h5 * P.dy_fac * yb.kt, sb);
    xb.t[0] += sb.tx * wo;
    xb.dn[TDT1] += sb.tx * wm;
    xb.t[2] += sb.tz * wo;
    xb.dn[TDT3] += sb.tz * wm;
    xb.t[1] += sb.ty;
    xb.c[G11] += dx2 * sb.k11 * wo;
    xb.dn[TD11] += dx2 * sb.k11 * wm;
    xb.dn[TD22] += dy2 * sb.k22 * wo;
    xb.dd += dy2 * sb.k22 * wm;
    xb.c[G33] += dz2 * sb.k33 * wo;
    xb.dn[TD33] += dz2 * sb.k33 * wm;
    xb.c[G12] += dy2 * sb.k12;
    xb.c[G21] += dx2 * sb.k12;
    xb.c[G13] += dz2 * sb.k13 * wo;
    xb.dn[TD13] += dz2 * sb.k13 * wm;
    xb.c[G31] += dx2 * sb.k13 * wo;
    xb.dn[TD31] += dx2 * sb.k13 * wm;
    xb.c[G23] += dz2 * sb.k23;
    xb.c[G32] += dy2 * sb.k23;
  }
}

}  // namespace lespt
