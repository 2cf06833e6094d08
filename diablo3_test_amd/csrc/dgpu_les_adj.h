// dgpu_les_adj.h -- adjoint kernels of the subgrid-model terms of the
// substep (FP64).  Conventions as in dgpu_adjoint.h: for a real loss L of a
// complex field c the adjoint holds dL/dRe c + i dL/dIm c, the transpose of
// multiplication by i k is multiplication by -i k, and the physical adjoints
// are plain gradients.  Every kernel is written in gather form (each output
// point sums the contributions that reach it), so no atomics are needed.
#pragma once
#include "dgpu_les.h"

namespace dgpu {

// ----------------------------------------------- stress divergence (T) --
// Transpose of k_les_div: from the adjoints of the subgrid forcings to those
// of the seven stress spectra and the two scalar flux spectra per scalar.
struct AdjLesDivArgs {
  const double2* f1;  // adjoint of lesf1..3, lesfth (the fbar of the RHS pass)
  const double2* f2;
  const double2* f3;
  const double2* fth[MAXTH];
  double2* t11;  // out
  double2* t33;
  double2* t13;
  double2* t12;
  double2* t23;
  double2* c1;
  double2* c3;
  double2* th_f1[MAXTH];
  double2* th_f3[MAXTH];
  const double* kx;
  const double* kz;
  const double* dyf;
  int js, j2, nth, th_sign;
  int jst[MAXTH], jet[MAXTH];
};

__global__ void __launch_bounds__(256) k_adj_les_div(AdjLesDivArgs a, Geo g) {
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
  auto in1 = [&](int jj) { return jj >= a.js && jj <= a.j2; };
  const bool i1 = in1(j), i1m = (j >= 1) && in1(j - 1), i2 = (j >= 2 && j <= g.ny);
  const C Z = {0, 0};
  const C f1 = i1 ? L(a.f1, o) : Z, f3 = i1 ? L(a.f3, o) : Z;
  const C f1m = i1m ? L(a.f1, o - sp) : Z, f3m = i1m ? L(a.f3, o - sp) : Z;
  const C f2 = i2 ? L(a.f2, o) : Z;
  const double r = i1 ? 1.0 / a.dyf[j] : 0.0, rm = i1m ? 1.0 / a.dyf[j - 1] : 0.0;
  C* w;
  w = reinterpret_cast<C*>(a.t11);
  w[o] = ik(kxq, f1);
  w = reinterpret_cast<C*>(a.t13);
  w[o] = ik(kzk, f1) + ik(kxq, f3);
  w = reinterpret_cast<C*>(a.t33);
  w[o] = ik(kzk, f3);
  w = reinterpret_cast<C*>(a.t12);
  w[o] = r * f1 - rm * f1m + ik(kxq, f2);
  w = reinterpret_cast<C*>(a.t23);
  w[o] = r * f3 - rm * f3m + ik(kzk, f2);
  w = reinterpret_cast<C*>(a.c1);
  w[o] = Z - ik(kxq, f2);
  w = reinterpret_cast<C*>(a.c3);
  w[o] = Z - ik(kzk, f2);
  for (int n = 0; n < a.nth; ++n) {
    const bool it = j >= a.jst[n] && j <= a.jet[n];
    const C ft = it ? L(a.fth[n], o) : Z;
    const double s = -(double)a.th_sign;
    reinterpret_cast<C*>(a.th_f1[n])[o] = s * ik(kxq, ft);
    reinterpret_cast<C*>(a.th_f3[n])[o] = s * ik(kzk, ft);
  }
}

// ----------------------------------------------------- velocity model --
// Gather of the velocity model's VJP: the gradient adjoints at plane j come
// from the model evaluations at planes j-1 (which read plane j as "above"),
// j, and j+1 (which read it as "below"); u1, u3 at plane j from the
// evaluations at j and j+1.  Gbar is written (the scalar model's adjoint,
// run first, is added), u1bar/u3bar are added to.
struct AdjLesVelArgs {
  LesVelArgs<double> fw;    // the forward inputs (gradients, u1, u3) and constants
  const double* bar[8];     // adjoints of tau11 tau33 tau13 tau12 tau23 nu_t cr1 cr3
  const double* gbar_in[9]; // gradient adjoints from the scalar model (or null)
  double* gbar[9];          // out
  double* u1bar;            // added to
  double* u3bar;
};

__global__ void __launch_bounds__(256) k_adj_les_vel(AdjLesVelArgs a, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int j = plane_of(idx, pp);
  if (j > g.ny + 1) return;
  const long long o = idx;
  double gb[9];
  for (int m = 0; m < 9; ++m) gb[m] = a.gbar_in[m] ? a.gbar_in[m][o] : 0.0;
  double u1b = 0.0, u3b = 0.0;
  const int upm[4] = {1, 3, 5, 7};
  const int dnm[5] = {0, 4, 8, 2, 6};
  for (int e = j - 1; e <= j + 1; ++e) {
    if (e < 1 || e > g.ny + 1) continue;
    const long long oe = o + (long long)(e - j) * pp;
    const lespt::VelPlane<double> P = les_vel_plane(a.fw, g, e);
    lespt::VelIn<double> x;
    les_vel_load(a.fw, g, e, oe, x);
    lespt::VelOut<double> yb;
    yb.t11 = a.bar[0][oe];
    yb.t33 = a.bar[1][oe];
    yb.t13 = a.bar[2][oe];
    yb.t12 = a.bar[3][oe];
    yb.t23 = a.bar[4][oe];
    yb.nt = a.bar[5][oe];
    yb.cr1 = a.bar[6][oe];
    yb.cr3 = a.bar[7][oe];
    lespt::VelIn<double> xb = {};
    lespt::vel_point_vjp(P, x, yb, xb);
    if (e == j) {
      for (int m = 0; m < 9; ++m) gb[m] += xb.c[m];
      u1b += xb.u1o;
      u3b += xb.u3o;
    } else if (e == j - 1) {
      for (int t = 0; t < 4; ++t) gb[upm[t]] += xb.up[t];
    } else {
      for (int t = 0; t < 5; ++t) gb[dnm[t]] += xb.dn[t];
      u1b += xb.u1m;
      u3b += xb.u3m;
    }
  }
  for (int m = 0; m < 9; ++m) a.gbar[m][o] = gb[m];
  a.u1bar[o] += u1b;
  a.u3bar[o] += u3b;
}

// ------------------------------------------------------- scalar model --
// Types 1/4: kappa_t = nu_t, f1 = kappa_t T1, f3 = kappa_t T3 pointwise.
// Type 5: gather over the evaluations at planes j-1, j, j+1, j+2.
struct AdjLesThArgs {
  LesThArgs<double> fw;
  const double* f1bar;     // adjoints of the outputs
  const double* f3bar;
  const double* ktbar;
  double* gbar[9];         // type 5: written (the first scalar) or added to
  int gbar_add;
  double* tbar[3];         // out: adjoints of T1 T2 T3
  double* ntbar;           // types 1/4: nu_t adjoint, added to
};

__global__ void __launch_bounds__(256) k_adj_les_th(AdjLesThArgs a, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int j = plane_of(idx, pp);
  if (j > g.ny + 1) return;
  const long long o = idx;
  if (a.fw.type != 5) {
    double f1b = 0, f3b = 0, kb = 0;
    if (j >= 1) {
      f1b = a.f1bar[o];
      f3b = a.f3bar[o];
      kb = a.ktbar[o];
    }
    const double kt = a.fw.nu_t[o];
    a.ntbar[o] += kb + f1b * a.fw.T1[o] + f3b * a.fw.T3[o];
    a.tbar[0][o] = f1b * kt;
    a.tbar[1][o] = 0.0;
    a.tbar[2][o] = f3b * kt;
    return;
  }
  double gb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double tb[3] = {0, 0, 0};
  const int upm[4] = {1, 3, 5, 7};
  const int dnm[5] = {0, 4, 8, 2, 6};
  for (int e = j - 1; e <= j + 2; ++e) {
    if (e < 1 || e > g.ny + 1) continue;
    const long long oe = o + (long long)(e - j) * pp;
    const lespt::ThPlane<double> P = les_th_plane(a.fw, g, e);
    if (!P.gyf_in && !P.gy_in) continue;
    lespt::ThIn<double> x;
    les_th_load(a.fw, g, e, oe, x);
    lespt::ThOut<double> yb = {a.f1bar[oe], a.f3bar[oe], a.ktbar[oe]};
    lespt::ThIn<double> xb = {};
    lespt::th_point_vjp(P, x, yb, xb);
    if (e == j) {
      for (int m = 0; m < 9; ++m) gb[m] += xb.c[m];
      for (int t = 0; t < 3; ++t) tb[t] += xb.t[t];
    } else if (e == j - 1) {
      for (int t = 0; t < 4; ++t) gb[upm[t]] += xb.up[t];
      tb[1] += xb.up[4];
    } else if (e == j + 1) {
      for (int t = 0; t < 5; ++t) gb[dnm[t]] += xb.dn[t];
      tb[0] += xb.dn[5];
      tb[2] += xb.dn[6];
    } else {
      gb[4] += xb.dd;
    }
  }
  for (int m = 0; m < 9; ++m) a.gbar[m][o] = (a.gbar_add ? a.gbar[m][o] : 0.0) + gb[m];
  for (int t = 0; t < 3; ++t) a.tbar[t][o] = tb[t];
}

// ------------------------------------------- derivative operators (T) --
// cbar += D^T X for the spectral derivative D of k_zinv_op (X the transposed
// inverse transform of a gradient adjoint): -i kx, -i kz, or the transposed
// two-point differences of OP_DY_GY / OP_DY_GYF.
__global__ void __launch_bounds__(256) k_adj_deriv(const double2* __restrict__ X, double2* __restrict__ cbar,
                                                   int op, const double* __restrict__ kx,
                                                   const double* __restrict__ kz, const double* __restrict__ dy,
                                                   const double* __restrict__ dyf, Geo g) {
  typedef cpx<double> C;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long sp = g.sp;
  const int j = plane_of(idx, sp);
  if (j >= g.ny2) return;
  const long long m = idx % sp;
  const int q = (int)((unsigned)m / (unsigned)g.nkk), kk = (int)((unsigned)m % (unsigned)g.nkk);
  const C* x = reinterpret_cast<const C*>(X);
  C v = {0, 0};
  switch (op) {
    case OP_DX: v = C{0, 0} - ik(kx[q], x[idx]); break;
    case OP_DZ: v = C{0, 0} - ik(kz[kk], x[idx]); break;
    case OP_DY_GY:  // G(j) = (c(j) - c(j-1)) / dy(j), j >= 1
      if (j >= 1) v = v + (1.0 / dy[j]) * x[idx];
      if (j + 1 < g.ny2) v = v - (1.0 / dy[j + 1]) * x[idx + sp];
      break;
    case OP_DY_GYF:  // G(j) = (c(j+1) - c(j)) / dyf(j), j <= Ny
      if (j >= 1 && j - 1 <= g.ny) v = v + (1.0 / dyf[j - 1]) * x[idx - sp];
      if (j <= g.ny) v = v - (1.0 / dyf[j]) * x[idx];
      break;
    default: v = x[idx]; break;
  }
  C* c = reinterpret_cast<C*>(cbar);
  c[idx] = c[idx] + v;
}

// ------------------------------------------------ CN solve with nu_t (T) --
// Adjoint of k_thomas_les (FP64, not the increment form) for one column per
// thread.  With A(nu) x = vec(r, u_old, nu): lambda = A^-T xbar, then
//   rbar = lambda on assembled rows,
//   ubar_old = (d vec / d u_old)^T lambda,
//   nubar += lambda^T (d vec / d nu - (d A / d nu) x).
// x is the forward solution over the whole column.  b', g' scratch as in
// the forward kernel.
__global__ void __launch_bounds__(256) k_adj_thomas_les(const double* __restrict__ xbar,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ uold,
                                                        const double* __restrict__ nt,
                                                        double* __restrict__ rbar, double* __restrict__ ubar,
                                                        double* __restrict__ ntbar, double* __restrict__ sb,
                                                        double* __restrict__ sg, const TriRow* __restrict__ rows,
                                                        int n, long long pp) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= pp) return;
  auto at = [&](const double* p, int j) { return p[(long long)j * pp + col]; };
  // lane rows: a(j), b(j), c(j) with the eddy viscosity on the row's faces
  auto faces = [&](int j, double& nl, double& nu) {
    const TriRow& rw = rows[j];
    if (rw.vtype) {
      nl = 0.5 * (at(nt, j) + at(nt, j - 1));
      nu = 0.5 * (at(nt, j) + at(nt, j + 1));
    } else {
      nl = at(nt, j);
      nu = at(nt, j + 1);
    }
  };
  auto abc = [&](int j, double& a, double& b, double& c) {
    const TriRow& rw = rows[j];
    a = rw.a;
    b = rw.b;
    c = rw.c;
    if (rw.kind == 1) {
      double nl, nu;
      faces(j, nl, nu);
      a -= rw.lm * nl;
      b += rw.lp * nu + rw.lm * nl;
      c -= rw.lp * nu;
    }
  };
  // forward elimination of A^T: sub(j) = c(j-1), diag b(j), sup(j) = a(j+1)
  double aj, bj, cj;
  abc(0, aj, bj, cj);
  double bp = bj, gp = at(xbar, 0), cprev = cj;
  sb[col] = bp;
  sg[col] = gp;
  for (int j = 1; j < n; ++j) {
    abc(j, aj, bj, cj);
    const double mlt = -cprev / bp;  // sub(j) / b'(j-1)
    // sup(j-1) = a(j)
    bp = bj + mlt * aj;
    gp = at(xbar, j) + mlt * gp;
    cprev = cj;
    sb[(long long)j * pp + col] = bp;
    sg[(long long)j * pp + col] = gp;
  }
  // back substitution: lambda(n-1) = g'/b', lambda(j) = (g'(j) - a(j+1) lambda(j+1)) / b'(j)
  double lam = gp / bp;
  rbar[(long long)(n - 1) * pp + col] = lam;  // provisional lambda store
  for (int j = n - 2; j >= 0; --j) {
    double a1, b1, c1;
    abc(j + 1, a1, b1, c1);
    lam = (sg[(long long)j * pp + col] - a1 * lam) / sb[(long long)j * pp + col];
    rbar[(long long)j * pp + col] = lam;
  }
  // rbar, ubar_old and nubar from lambda (now in rbar)
  for (int j = 0; j < n; ++j) ubar[(long long)j * pp + col] = 0.0;
  for (int j = 0; j < n; ++j) {
    const TriRow& rw = rows[j];
    const long long o = (long long)j * pp + col;
    const double l = rbar[o];
    if (rw.kind != 1) {
      rbar[o] = 0.0;
      continue;
    }
    double nl, nu;
    faces(j, nl, nu);
    const double um = at(uold, j - 1), uc = at(uold, j), up = at(uold, j + 1);
    const double xm = at(x, j - 1), xc = at(x, j), xp = at(x, j + 1);
    // vec(j) = r + wp (up - uc) - wm (uc - um) + lp nu (up - uc) - lm nl (uc - um)
    ubar[o + pp] += l * (rw.wp + rw.lp * nu);
    ubar[o] += l * (-rw.wp - rw.wm - rw.lp * nu - rw.lm * nl);
    ubar[o - pp] += l * (rw.wm + rw.lm * nl);
    const double nub = l * rw.lp * ((up - uc) + (xp - xc));
    const double nlb = -l * rw.lm * ((uc - um) + (xc - xm));
    if (rw.vtype) {
      ntbar[o] += 0.5 * (nlb + nub);
      ntbar[o - pp] += 0.5 * nlb;
      ntbar[o + pp] += 0.5 * nub;
    } else {
      ntbar[o] += nlb;
      ntbar[o + pp] += nub;
    }
  }
}

// ------------------------------------------------------ wall BCs (T) ----
// Transposes of k_bc_post followed by k_les_bc (the wall and ghost values
// written at the start of the LES substep), applied in reverse order on the
// adjoint of the post-BC state.  Each forward assignment dst = s * src + k
// becomes src_bar += s * dst_bar, dst_bar = 0.
struct AdjBcArgs {
  double2* c1;
  double2* c2;
  double2* c3;
  double2* cth[MAXTH];
  int nth;
  int u_lo, v_lo, w_lo, u_hi, v_hi, w_hi;
  int th_lo[MAXTH], th_hi[MAXTH];
};

__global__ void k_adj_les_bc(AdjBcArgs a, Geo g) {
  typedef cpx<double> C;
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.sp) return;
  const bool mean = (m == 0);
  const long long P = g.sp;
  const long long T0 = (long long)g.ny * P + m, T1 = T0 + P, Tm = T0 - P;
  const long long B0 = m, B1 = P + m, B2 = 2 * P + m;
  const C Z = {0, 0};
  // dst <- src (copy): src += dst, dst = 0
  auto undo_copy = [](C* c, long long dst, long long src, double s) {
    c[src] = c[src] + s * c[dst];
    c[dst] = C{0, 0};
  };
  C* u1 = reinterpret_cast<C*>(a.c1);
  C* u2 = reinterpret_cast<C*>(a.c2);
  C* u3 = reinterpret_cast<C*>(a.c3);
  // ---- k_les_bc, in reverse
  if (a.w_lo == 1) {
    undo_copy(u3, B0, B1, 1.0);
    undo_copy(u3, B1, B2, 1.0);
  }
  if (a.u_lo == 1) {
    undo_copy(u1, B0, B1, 1.0);
    undo_copy(u1, B1, B2, 1.0);
  }
  if (a.w_hi == 1) {
    undo_copy(u3, T1, T0, 1.0);
    undo_copy(u3, T0, Tm, 1.0);
  }
  if (a.u_hi == 1) {
    undo_copy(u1, T1, T0, 1.0);
    undo_copy(u1, T0, Tm, 1.0);
  }
  for (int n = a.nth - 1; n >= 0; --n) {
    C* c = reinterpret_cast<C*>(a.cth[n]);
    if (a.th_lo[n] == 1) {
      undo_copy(c, B0, B1, 1.0);
      undo_copy(c, B1, B2, 1.0);
    }
    if (a.th_hi[n] == 1) {
      undo_copy(c, T1, T0, 1.0);
      undo_copy(c, T0, Tm, 1.0);
    }
    // boundary.f90:728-821 part of k_les_bc
    if (a.th_hi[n] == 0) {
      c[T0] = Z;
      c[T1] = Z;
    } else {
      undo_copy(c, T1, T0, 1.0);
      undo_copy(c, T0, Tm, 1.0);
    }
    if (a.th_lo[n] == 0) {
      c[B0] = Z;
      c[B1] = Z;
    } else {
      undo_copy(c, B0, B1, 1.0);
      undo_copy(c, B1, B2, 1.0);
    }
  }
  // ---- k_bc_post, in reverse (upper wall, then lower)
  if (a.v_hi == 0) {
    if (mean) undo_copy(u2, T1, T0, -1.0);
  } else {
    undo_copy(u2, T1, T0, 1.0);
    undo_copy(u2, T0, Tm, 1.0);
  }
  if (a.w_hi == 0) {
    u3[T0] = Z;
    u3[T1] = Z;
  } else {
    undo_copy(u3, T1, T0, 1.0);
    undo_copy(u3, T0, Tm, 1.0);
  }
  if (a.u_hi == 0) {
    u1[T0] = Z;
    u1[T1] = Z;
  } else {
    undo_copy(u1, T1, T0, 1.0);
    undo_copy(u1, T0, Tm, 1.0);
  }
  if (a.v_lo == 0) {
    undo_copy(u2, B0, B1, 1.0);
    undo_copy(u2, B1, B2, -1.0);
  } else {
    undo_copy(u2, B0, B1, 1.0);
    undo_copy(u2, B1, B2, 1.0);
  }
  if (a.w_lo == 0) {
    u3[B0] = Z;
    u3[B1] = Z;
  } else {
    undo_copy(u3, B0, B1, 1.0);
    undo_copy(u3, B1, B2, 1.0);
  }
  if (a.u_lo == 0) {
    u1[B0] = Z;
    u1[B1] = Z;
  } else {
    undo_copy(u1, B0, B1, 1.0);
    undo_copy(u1, B1, B2, 1.0);
  }
}

// y += x over n complex values
__global__ void __launch_bounds__(256) k_axpy2(double2* __restrict__ y, const double2* __restrict__ x, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    y[i].x += x[i].x;
    y[i].y += x[i].y;
  }
}

}  // namespace dgpu
