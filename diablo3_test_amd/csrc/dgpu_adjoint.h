// dgpu_adjoint.h -- adjoint (vector-Jacobian product) of the FP64 DNS
// substep rk_chan_1 (for/channel.f90:10-917) and rem_div_chan (:946-1023).
//
// Conventions.  Complex arrays are real pairs; the adjoint of a complex array
// c is lam = dL/dRe c + i dL/dIm c.  A complex-linear map y = m x then has the
// adjoint xbar += conj(m) ybar; ik products become -ik.  The transforms
// (fft.f90:247-382) as real-linear maps have the transposes
//   (forward F)^T lam = P(W lam) / (Nx Nz),     W = 1 at kx = 0, 1/2 at kx > 0,
//   (inverse P)^T ubar = Nx Nz W' F(ubar),       W' = 1 at kx = 0, 2 at kx > 0,
// so the adjoint reuses the forward passes with per-kx scale factors.  The
// Thomas solves become solves with the transposed tridiagonal matrices
// ("reverse" sweeps), and the products' Jacobians are gathered over the
// three-plane stencils of channel.f90:299-459, 557-608.
#pragma once
#include "dgpu_kernels.h"

namespace dgpu {

// ------------------------------------------------- projection adjoint ----
// rem_div_chan's map (cu1, cu2, cu3, cp) -> (cu1 - ikx phi, cu2 - d phi/dy,
// cu3 - ikz phi, cp + phi/h) with phi = T^-1 div(cu), per (kx, kz) mode:
// phibar from the outputs, y = T^-T phibar, then the divergence's transpose.
struct AdjRemDivArgs {
  double2* a1;   // in/out: adjoints of cu1, cu2, cu3
  double2* a2;
  double2* a3;
  const double2* ap;  // adjoint of cp (the map passes it through)
  const double* ib;   // 1/b_j of the forward matrix (k_remdiv_pivots)
  const double* kx;
  const double* kz;
  double inv_h;
  int js, je;
};

// Transpose of k_remdiv_seg.  With phibar the adjoint of phi (gathered from
// the outputs cu -= grad phi and cp += phi/h), ybar = T^-T phibar by the
// factors of the forward matrix T = L U (pivots b_j, L's multipliers
// ml_j / b_{j-1}): T^T = U^T L^T, so
//   z_j = (phibar_j - mu_{j-1} z_{j-1}) / b_j                  (forward, U^T)
//   y_j = z_j - (ml_{j+1} / b_j) y_{j+1}                        (backward, L^T)
// and the divergence rows 2..Ny-1 take their transpose: cu1bar -= ikx y,
// cu3bar -= ikz y, cu2bar(j+1) += y/dyf(j), cu2bar(j) -= y/dyf(j).  Row 1 is
// no divergence row, so its y is not needed; its z_1 = phibar_1 (b_1 = 1)
// enters segment 0 as the forward carry.  Layout, segments and scans as in
// k_remdiv_seg.
template <int L, int MW, int MAXT>
__global__ void __launch_bounds__(MAXT) k_adj_remdiv_seg(AdjRemDivArgs a, Geo g, const PoisRow* __restrict__ rows,
                                                         int S) {
  typedef cpx<double> C;
  constexpr int SPW = 64 / MW;
  constexpr int SMAX = MAXT / MW, WMAX = MAXT / 64;
  __shared__ double tP[WMAX][MW];
  __shared__ C tA[WMAX][MW];
  __shared__ double tQ[WMAX][MW];
  __shared__ C tB[WMAX][MW];
  __shared__ C tL[SMAX][MW];
  __shared__ C z1s[MW];
  // the row coefficients staged in LDS once, as in k_remdiv_seg
  __shared__ PoisRow srw[2 + L * SMAX];
  for (int i = threadIdx.x; i <= g.ny && i < 2 + L * SMAX; i += blockDim.x) srw[i] = rows[i];
  __syncthreads();
  const int mm = threadIdx.x % MW, s = threadIdx.x / MW;
  const int w = threadIdx.x / 64, sl = (threadIdx.x % 64) / MW;
  const int nw = blockDim.x / 64;
  const long long sp = g.sp;
  const long long m = (long long)xcd_block(blockIdx.x, gridDim.x) * MW + mm;  // as k_remdiv_seg
  const bool live = m < sp;
  const long long mc = live ? m : 0;
  const int q = (int)((unsigned)mc / (unsigned)g.nkk), kk = (int)((unsigned)mc % (unsigned)g.nkk);
  const double kxq = a.kx[q], kzk = a.kz[kk];
  const double mu1 = (mc == 0) ? 0.0 : -1.0;
  const int ny = g.ny;
  const int j0 = 2 + s * L;
  const C Z = {0.0, 0.0};
  auto ldc = [](const double2* base, unsigned ob) {
    const double2 v = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + ob);
    return C{v.x, v.y};
  };
  auto stc = [](double2* base, unsigned ob, C v) {
    *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + ob) = make_double2(v.x, v.y);
  };
  auto ldd = [](const double* base, unsigned ob) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + ob);
  };
  auto off = [&](int j) { return (unsigned)(((long long)j * sp + m) * 16); };
  const long long hp = piv_plane(g), mh = piv_index(g, q, kk);  // pivots shared by +-kz
  auto offp = [&](int j) { return (unsigned)(((long long)j * hp + mh) * 8); };
  auto shup = [](C v, int d) { return C{__shfl_up(v.x, d), __shfl_up(v.y, d)}; };
  auto shdn = [](C v, int d) { return C{__shfl_down(v.x, d), __shfl_down(v.y, d)}; };
  auto inp = [&](int j) { return j >= a.js && j <= a.je; };
  // adjoints at rows j0 .. j0+L-1 (plus cu2bar at j0+L) and 1/b
  C d1[L], d3[L], d2[L + 1], pb[L];
  double ibv[L];
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    const unsigned o = off(j);
    const bool ok = live && j <= ny;
    d1[t] = ok ? ldc(a.a1, o) : Z;
    d3[t] = ok ? ldc(a.a3, o) : Z;
    d2[t] = ok ? ldc(a.a2, o) : Z;
    pb[t] = (ok && inp(j)) ? ldc(a.ap, o) : Z;
    ibv[t] = ok ? ldd(a.ib, offp(j)) : 0.0;
  }
  d2[L] = (live && j0 + L <= ny) ? ldc(a.a2, off(j0 + L)) : Z;
  // phibar of each row
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    if (j > ny) break;
    C v = ik(kxq, d1[t]) + ik(kzk, d3[t]) - srw[j].idy * d2[t];
    if (j + 1 <= ny) v = v + srw[j + 1].idy * d2[t + 1];
    if (inp(j)) v = v + a.inv_h * pb[t];
    pb[t] = v;
  }
  if (s == 0) {  // z_1 = phibar_1 (row 1: no cu2bar(1) term)
    C z1 = Z;
    if (live) {
      const unsigned o = off(1);
      z1 = ik(kxq, ldc(a.a1, o)) + ik(kzk, ldc(a.a3, o));
      if (ny >= 2) z1 = z1 + srw[2].idy * d2[0];
      if (inp(1)) z1 = z1 + a.inv_h * ldc(a.ap, o);
    }
    z1s[mm] = z1;
  }
  // forward sweep (U^T) of the segment from a zero carry: z_t = gl_t + Pl_t * z(j0-1)
  C gl[L];
  double Pl[L];
  {
    C h = Z;
    double P = 1.0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int j = j0 + t;
      if (j <= ny) {
        const double mup = (j == 2) ? mu1 : srw[j - 1].mu;
        const double f = -mup * ibv[t];
        h = ibv[t] * pb[t] + f * h;
        P = f * P;
      }
      gl[t] = h;
      Pl[t] = P;
    }
  }
  {
    double P = Pl[L - 1];
    C A = gl[L - 1];
#pragma unroll
    for (int d = 1; d < SPW; d <<= 1) {
      const double Pp = __shfl_up(P, d * MW);
      const C Ap = shup(A, d * MW);
      if (sl >= d) {
        A = A + P * Ap;
        P = P * Pp;
      }
    }
    if (sl == SPW - 1) {
      tP[w][mm] = P;
      tA[w][mm] = A;
    }
    double Pe = __shfl_up(P, MW);
    C Ae = shup(A, MW);
    if (sl == 0) {
      Pe = 1.0;
      Ae = Z;
    }
    __syncthreads();
    C cin = z1s[mm];  // z at row 1 enters wavefront 0
    for (int w2 = 0; w2 < w; ++w2) cin = tA[w2][mm] + tP[w2][mm] * cin;
    const C zin = Ae + Pe * cin;
#pragma unroll
    for (int t = 0; t < L; ++t) gl[t] = gl[t] + Pl[t] * zin;
  }
  // backward sweep (L^T) of the segment from a zero carry: y_t = gl_t + Pl_t * y(j0+L)
  {
    C x = Z;
    double Q = 1.0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
      const int j = j0 + t;
      if (j <= ny) {
        const double f = (j < ny) ? -srw[j + 1].ml * ibv[t] : 0.0;
        x = gl[t] + f * x;
        Q = f * Q;
      }
      gl[t] = x;
      Pl[t] = Q;
    }
  }
  {
    double Q = Pl[0];
    C B = gl[0];
#pragma unroll
    for (int d = 1; d < SPW; d <<= 1) {
      const double Qn = __shfl_down(Q, d * MW);
      const C Bn = shdn(B, d * MW);
      if (sl + d < SPW) {
        B = B + Q * Bn;
        Q = Q * Qn;
      }
    }
    if (sl == 0) {
      tQ[w][mm] = Q;
      tB[w][mm] = B;
    }
    double Qe = __shfl_down(Q, MW);
    C Be = shdn(B, MW);
    if (sl == SPW - 1) {
      Qe = 1.0;
      Be = Z;
    }
    __syncthreads();
    C ctop = Z;
    for (int w2 = nw - 1; w2 > w; --w2) ctop = tB[w2][mm] + tQ[w2][mm] * ctop;
    const C ytop = Be + Qe * ctop;
#pragma unroll
    for (int t = 0; t < L; ++t) gl[t] = gl[t] + Pl[t] * ytop;  // gl now holds y
  }
  // y at the row below the segment (row 1 for segment 0, not a divergence row)
  tL[s][mm] = gl[L - 1];
  __syncthreads();
  const C ym = (s > 0) ? tL[s - 1][mm] : Z;
  if (!live) return;
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    if (j > ny) break;
    const unsigned o = off(j);
    const C y = gl[t];
    const bool dr = j <= ny - 1;                 // j >= 2 holds for every segment row
    const bool dm = j - 1 >= 2 && j - 1 <= ny - 1;
    const C yprev = (t == 0) ? ym : gl[t - 1];
    C d = d2[t];
    if (dm) d = d + srw[j - 1].idyf * yprev;
    if (dr) d = d - srw[j].idyf * y;
    stc(a.a2, o, d);
    if (dr) {
      stc(a.a1, o, d1[t] - ik(kxq, y));
      stc(a.a3, o, d3[t] - ik(kzk, y));
    }
  }
}

// ------------------------------------------------------ CN adjoint ----
// The transposed solve vbar = A^-T xbar and its epilogue (rbar = vbar on the
// assembled rows, ubar = E^T vbar) are k_thomas_seg with ADJ set.

// ------------------------------------------------- spectral RHS adjoint ----
// Part a (per spectral entry): fbar = cfbar_new + temp5 crbar on the rows
// where the forward wrote cf (0 elsewhere), and cfbar_old = temp3 crbar there
// (rk_step > 1; cf_old is read only through cr), cfbar_new elsewhere.
struct AdjRhsArgs {
  const double2* acr[3 + MAXTH];  // crbar: cr1, cr2, cr3, crth_n (natural order)
  double2* acf[3 + MAXTH];        // in: cfbar of the output; out: cfbar of the input
  double2* fbar[3 + MAXTH];       // out (part a) / in (part b)
  double2* acu[3];                // in/out: cubar
  double2* acp;                   // in/out: cpbar
  double2* acth[MAXTH];           // in/out: cthbar
  const double* kx;
  const double* kz;
  const double* dy;
  const double* dyf;
  double nu, beta, coriolis;
  double grav_x, grav_y, grav_z;
  double pr[MAXTH], dthdx[MAXTH], dthdz[MAXTH];
  double temp3, temp4, temp5;
  // Front thermal-wind coefficients, as in RhsArgs (null unless Front)
  const double* fr_yf;
  const double* fr_y;
  double fr_sx, fr_sz, fr_dr;
  int rk_step, nth;
  int js, je;
  int jst[MAXTH], jet[MAXTH];
};

__device__ __forceinline__ bool adj_in_range(const AdjRhsArgs& a, const Geo& g, int c, int j) {
  if (c == 0 || c == 2) return j >= a.js && j <= a.je;
  if (c == 1) return j >= 2 && j <= g.ny;
  return j >= a.jst[c - 3] && j <= a.jet[c - 3];
}

__global__ void __launch_bounds__(256) k_adj_rhs_a(AdjRhsArgs a, Geo g) {
  typedef cpx<double> C;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.sp * g.ny2) return;
  const int j = plane_of(idx, g.sp);
  const int nc = 3 + a.nth;
  for (int c = 0; c < nc; ++c) {
    C* cf = reinterpret_cast<C*>(a.acf[c]);
    C* fb = reinterpret_cast<C*>(a.fbar[c]);
    if (adj_in_range(a, g, c, j)) {
      const C cr = reinterpret_cast<const C*>(a.acr[c])[idx];
      fb[idx] = cf[idx] + a.temp5 * cr;
      cf[idx] = (a.rk_step > 1) ? a.temp3 * cr : C{0.0, 0.0};
    } else {
      fb[idx] = C{0.0, 0.0};
    }
  }
}

// Part b: the linear explicit terms of channel.f90:39-253 read cu1..3, cp and
// th; their transposes, gathered per entry (the buoyancy and -temp4 dcp/dy
// terms couple planes j and j+1).
__global__ void __launch_bounds__(256) k_adj_rhs_b(AdjRhsArgs a, Geo g) {
  typedef cpx<double> C;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.sp * g.ny2) return;
  const int j = plane_of(idx, g.sp);
  const long long mm = (long long)((unsigned)idx % (unsigned)g.sp);
  const int q = (int)((unsigned)mm / (unsigned)g.nkk), kk = (int)((unsigned)mm % (unsigned)g.nkk);
  const double kxq = a.kx[q], kzk = a.kz[kk];
  const double lap = kxq * kxq + kzk * kzk;
  const double visc = (a.beta == 1.0) ? lap : pow(lap, a.beta);
  const bool mean = (q == 0 && kk == 0);
  const long long sp = g.sp;
  const int ny = g.ny;
  auto L = [&](const double2* p, long long o) { return reinterpret_cast<const C*>(p)[o]; };
  auto inr = [&](int c, int jj) { return jj >= 0 && jj < g.ny2 && adj_in_range(a, g, c, jj); };
  const C zero = {0.0, 0.0};
  const C f1 = L(a.fbar[0], idx), f2 = L(a.fbar[1], idx), f3 = L(a.fbar[2], idx);
  const C r1 = inr(0, j) ? L(a.acr[0], idx) : zero;
  const C r2 = inr(1, j) ? L(a.acr[1], idx) : zero;
  const C r3 = inr(2, j) ? L(a.acr[2], idx) : zero;
  C b1 = r1 + (-a.nu * visc) * f1 - a.coriolis * f3;
  C b3 = r3 + (-a.nu * visc) * f3 + a.coriolis * f1;
  C b2 = r2 + (-a.nu * visc) * f2;
  // cr1 -= temp4 ikx cp, cr3 -= temp4 ikz cp, cr2 -= temp4 (cp(j) - cp(j-1))/dy(j)
  C bp = a.temp4 * (ik(kxq, r1) + ik(kzk, r3));
  bp = bp - (a.temp4 / a.dy[j > 0 ? j : 1]) * r2;
  if (j + 1 <= ny && inr(1, j + 1)) bp = bp + (a.temp4 / a.dy[j + 1]) * L(a.acr[1], idx + sp);
  const C f2up = (j + 1 < g.ny2) ? L(a.fbar[1], idx + sp) : zero;
  // The Front thermal wind (forcing.f90:50-100).  A term alpha i k x of the
  // forward adds -alpha i k xbar to the adjoint of x; the fbar are zero
  // outside their fields' row ranges.
  if (a.fr_yf) {
    const double cy = a.fr_yf[j];
    b1 = b1 + (a.fr_sx * cy) * ik(kzk, f1) - (a.fr_sz * cy) * ik(kxq, f1);
    b3 = b3 + (a.fr_sx * cy) * ik(kzk, f3) - (a.fr_sz * cy) * ik(kxq, f3);
    // f1(j) += sz dr (u2(j) + u2(j+1))/2, f3(j) -= sx dr (u2(j) + u2(j+1))/2
    C f1s = f1, f3s = f3;
    if (j > 0) {
      f1s = f1s + L(a.fbar[0], idx - sp);
      f3s = f3s + L(a.fbar[2], idx - sp);
    }
    b2 = b2 + (0.5 * a.fr_sz * a.fr_dr) * f1s - (0.5 * a.fr_sx * a.fr_dr) * f3s;
    const double cyy = a.fr_y[j];
    b2 = b2 + (a.fr_sx * cyy) * ik(kzk, f2) - (a.fr_sz * cyy) * ik(kxq, f2);
  }
  for (int n = 0; n < a.nth; ++n) {
    const C ft = L(a.fbar[3 + n], idx);
    const C rt = inr(3 + n, j) ? L(a.acr[3 + n], idx) : zero;
    b1 = b1 - a.dthdx[n] * ft;
    b3 = b3 - a.dthdz[n] * ft;
    C bt = rt + (-(a.nu / a.pr[n]) * visc) * ft + a.grav_x * f1 + a.grav_z * f3;
    if (a.fr_yf) {
      const double cy = a.fr_yf[j];
      bt = bt + (a.dthdx[n] * cy) * ik(kzk, ft) - (a.dthdz[n] * cy) * ik(kxq, ft);
    }
    if (!mean) {
      // f2(j) += grav_y (dyf(j-1) th(j) + dyf(j) th(j-1)) / (2 dy(j)) on rows 2..Ny
      if (j >= 2 && j <= ny) bt = bt + (a.grav_y * a.dyf[j - 1] / (2.0 * a.dy[j])) * f2;
      if (j + 1 >= 2 && j + 1 <= ny) bt = bt + (a.grav_y * a.dyf[j + 1] / (2.0 * a.dy[j + 1])) * f2up;
    }
    C* th = reinterpret_cast<C*>(a.acth[n]);
    th[idx] = th[idx] + bt;
  }
  C* u1 = reinterpret_cast<C*>(a.acu[0]);
  C* u2 = reinterpret_cast<C*>(a.acu[1]);
  C* u3 = reinterpret_cast<C*>(a.acu[2]);
  C* cp = reinterpret_cast<C*>(a.acp);
  u1[idx] = u1[idx] + b1;
  u2[idx] = u2[idx] + b2;
  u3[idx] = u3[idx] + b3;
  cp[idx] = cp[idx] + bp;
}

// -------------------------------------------- products adjoint, x pass ----
// From the x-spectra adjoints of the product slots to the physical adjoints
// of the products.  A slot X = ikx Fx(A) + Fx(B) gives Abar = Fx^T(-ikx Xbar),
// Bbar = Fx^T(Xbar); Fx^T lam = c2r(W lam) / Nx.  Each pair's (Abar, Bbar)
// comes out of one complex transform (real part Abar, imaginary part Bbar).
//   pair 0: (-ikx X0, X0)          -> u1u1, V1
//   pair 1: (X1 - ikx X2, X2)      -> u3u1, V3
//   pair 2: (X3, X5)               -> u3u3, W3
//   pair 3: (-ikx X4, X4)          -> W1, V2
//   pair 4+2n: (-ikx X6, X6)       -> th u1, Vt
//   pair 5+2n: (X7, 0)             -> th u3
struct AdjPairArgs {
  const double2* X[MAXSLOT];  // slot adjoints (x-spectra)
  double* outA[4 + 2 * MAXTH];
  double* outB[4 + 2 * MAXTH];
  const double* kx;
  int npairs;
};

// RAG: ragged last tile (Nz not a multiple of KT); rows k >= Nz are skipped.
// LOGN: the transform length at compile time, as in fft_any (0: Nx from g).
template <bool WAVE, bool RAG = false, int LOGN = 0>
__global__ void __launch_bounds__(512) k_adj_xpairs(AdjPairArgs a, Geo g, const cpx<double>* __restrict__ tw,
                                                    int j0) {
  typedef cpx<double> C;
  constexpr int KT = XTile<double>::KT;  // z lines per workgroup
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  const int G = g.nx >> 3;
  const int L = x_group_stride(g.nx);
  const int ngroups = blockDim.x / G;  // a power of two dividing KT
  const int lg_ng = 31 - __builtin_clz(ngroups);
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  C* fbuf = smem + grp * L;
  // the power-formed twiddles read one table entry a step, from the caches:
  // without the table's 8 KB two 512-thread workgroups share a CU (its 106
  // VGPRs allow them), and the adjoint's kernel class went 22.1 -> 19.7 ms per
  // step at 512x257x512
  const C* twl = LOGN == FFT_XL512P ? tw : lds_twiddles(smem + ngroups * L, tw, g.nx);
  const int j = j0 + blockIdx.y, k0 = blockIdx.x * KT;
  const double inv = 1.0 / g.nx;
  for (int pr = 0; pr < a.npairs; ++pr) {
    const int base = (pr >= 4) ? 6 + 2 * ((pr - 4) >> 1) : 0;
    const int kind = pr < 4 ? pr : 4 + ((pr - 4) & 1);
    for (int m0 = 0; m0 < KT; m0 += ngroups) {
      __syncthreads();  // the previous round's output reads of the buffers are done
      for (int qq = g.nkx + 1 + gt; qq < g.nx - g.nkx; qq += G) fbuf[pidx(qq)] = C{0.0, 0.0};
      // consecutive threads take consecutive z lines of one kx, so a
      // wavefront reads 16-byte values of ngroups lines as one run
      for (int idx = threadIdx.x; idx < g.nq * ngroups; idx += blockDim.x) {
        const int q = idx >> lg_ng, mm = idx & (ngroups - 1);
        const bool ok = !RAG || k0 + m0 + mm < g.nz;
        const long long o = (long long)j * g.tp + (long long)q * g.nz + (ok ? k0 + m0 + mm : 0);
        auto X = [&](int s) { return ok ? reinterpret_cast<const C*>(a.X[s])[o] : C{0.0, 0.0}; };
        const double kxq = a.kx[q];
        C al, be;
        switch (kind) {
          case 0: al = ik(-kxq, X(0)); be = X(0); break;
          case 1: al = X(1) + ik(-kxq, X(2)); be = X(2); break;
          case 2: al = X(3); be = X(5); break;
          case 3: al = ik(-kxq, X(4)); be = X(4); break;
          case 4: al = ik(-kxq, X(base)); be = X(base); break;
          default: al = X(base + 1); be = C{0.0, 0.0}; break;
        }
        const double w = (q == 0 ? 1.0 : 0.5) * inv;
        const C A = w * al, B = w * be;
        C* f = smem + mm * L;
        if (q == 0) {
          f[pidx(0)] = C{A.x, B.x};
        } else {
          f[pidx(q)] = C{A.x - B.y, A.y + B.x};
          f[pidx(g.nx - q)] = C{A.x + B.y, B.x - A.y};
        }
      }
      __syncthreads();  // every group's spectrum is in place
      C v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = fbuf[pidx(gt + G * r)];
      fft_any<double, 1, WAVE, LOGN>(v, fbuf, g.nx, g.lognx, twl, gt);
      if (RAG && k0 + m0 + grp >= g.nz) continue;  // past Nz: nothing to store (the barrier is at the loop top)
      const long long ro = (long long)j * g.pp + (long long)(k0 + m0 + grp) * g.nx;
      double* oa = a.outA[pr] + ro;
      double* ob = a.outB[pr] + ro;
      for (int i = gt; i < g.nx; i += G) {
        const C z = fbuf[pidx(i)];
        oa[i] = z.x;
        ob[i] = z.y;
      }
    }
  }
}

// ------------------------------------------ products adjoint, gather ----
// ubar(j) (+)= sum of dP/du(j) Pbar over the products P at rows j-1, j, j+1
// (channel.f90:299-459, 557-608); Pbar is zero on rows the forward skipped.
struct AdjProdArgs {
  const double* u1;
  const double* u2;
  const double* u3;
  const double* th[MAXTH];
  const double* A[4 + 2 * MAXTH];  // pair order of k_adj_xpairs
  const double* B[4 + 2 * MAXTH];
  double* ub1;
  double* ub2;
  double* ub3;
  double* ubth[MAXTH];
  const double* dy;
  const double* dyf;
  int nth;
  int store;  // 1: ubar = the gathered sum (nothing else has written it); 0: ubar +=
};

__global__ void __launch_bounds__(256) k_adj_prod(AdjProdArgs a, Geo g) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long pp = g.pp;
  const int n = g.ny2;
  if (idx >= pp * n) return;
  const int j = plane_of(idx, pp);
  const long long o = idx, up = idx + pp, dn = idx - pp;
  const bool hu = j + 1 < n, hd = j >= 1;
  auto V = [&](const double* f, long long i, bool ok) { return ok ? f[i] : 0.0; };
  const double u1c = a.u1[o], u2c = a.u2[o], u3c = a.u3[o];
  const double u1p = V(a.u1, up, hu), u2p = V(a.u2, up, hu), u3p = V(a.u3, up, hu);
  const double u1m = V(a.u1, dn, hd), u2m = V(a.u2, dn, hd), u3m = V(a.u3, dn, hd);
  auto c_of = [&](int jj) { return 1.0 / (2.0 * a.dyf[jj]); };
  const double cj = c_of(j), cm = hd ? c_of(j - 1) : 0.0, cp = hu ? c_of(j + 1) : 0.0;
  // pair Pbar values at j-1, j, j+1
  auto A = [&](int pr, long long i, bool ok) { return ok ? a.A[pr][i] : 0.0; };
  auto B = [&](int pr, long long i, bool ok) { return ok ? a.B[pr][i] : 0.0; };
  double b1 = 0.0, b2 = 0.0, b3 = 0.0;
  // pair 0: A0 = u1 u1; B0 = c (u1p u2p + u1 u2p - u1 u2 - u1m u2)
  b1 += 2.0 * u1c * A(0, o, true);
  b1 += cj * (u2p - u2c) * B(0, o, true) + cm * u2c * B(0, dn, hd) - cp * u2p * B(0, up, hu);
  b2 += -cj * (u1c + u1m) * B(0, o, true) + cm * (u1c + u1m) * B(0, dn, hd);
  // pair 1: A1 = u3 u1; B1 as B0 with u3
  b1 += u3c * A(1, o, true);
  b3 += u1c * A(1, o, true);
  b3 += cj * (u2p - u2c) * B(1, o, true) + cm * u2c * B(1, dn, hd) - cp * u2p * B(1, up, hu);
  b2 += -cj * (u3c + u3m) * B(1, o, true) + cm * (u3c + u3m) * B(1, dn, hd);
  // pair 2: A2 = u3 u3; B2 = I3 u2, I3(j) = (dyf(j) u3(j) + dyf(j-1) u3(j-1)) / (2 dy(j))
  b3 += 2.0 * u3c * A(2, o, true);
  {
    const double i3 = hd ? (a.dyf[j] * u3c + a.dyf[j - 1] * u3m) / (2.0 * a.dy[j]) : 0.0;
    b2 += i3 * B(2, o, true);
    b3 += (hd ? a.dyf[j] / (2.0 * a.dy[j]) : 0.0) * u2c * B(2, o, true);
    if (hu) b3 += (a.dyf[j] / (2.0 * a.dy[j + 1])) * u2p * B(2, up, true);
  }
  // pair 3: A3 = I1 u2; B3 = (1/4 (u2 + u2p)^2 - 1/4 (u2 + u2m)^2) / dy
  {
    const double i1 = hd ? (a.dyf[j] * u1c + a.dyf[j - 1] * u1m) / (2.0 * a.dy[j]) : 0.0;
    b2 += i1 * A(3, o, true);
    b1 += (hd ? a.dyf[j] / (2.0 * a.dy[j]) : 0.0) * u2c * A(3, o, true);
    if (hu) b1 += (a.dyf[j] / (2.0 * a.dy[j + 1])) * u2p * A(3, up, true);
    if (hd) b2 += (0.5 * (u2p - u2m) / a.dy[j]) * B(3, o, true);
    if (hd) b2 += (0.5 * (u2m + u2c) / a.dy[j - 1 > 0 ? j - 1 : 1]) * B(3, dn, true);
    if (hu) b2 += (-0.5 * (u2p + u2c) / a.dy[j + 1]) * B(3, up, true);
  }
  // scalars: A4 = th u1, B4 = c (thp u2p + th u2p - th u2 - thm u2), A5 = th u3
  for (int s = 0; s < a.nth; ++s) {
    const double* th = a.th[s];
    const double tc = th[o], tp = V(th, up, hu), tm = V(th, dn, hd);
    const int p4 = 4 + 2 * s, p5 = p4 + 1;
    double bt = 0.0;
    b1 += tc * A(p4, o, true);
    bt += u1c * A(p4, o, true);
    bt += cj * (u2p - u2c) * B(p4, o, true) + cm * u2c * B(p4, dn, hd) - cp * u2p * B(p4, up, hu);
    b2 += -cj * (tc + tm) * B(p4, o, true) + cm * (tc + tm) * B(p4, dn, hd);
    b3 += tc * A(p5, o, true);
    bt += u3c * A(p5, o, true);
    a.ubth[s][o] = a.store ? bt : a.ubth[s][o] + bt;
  }
  if (a.store) {
    a.ub1[o] = b1;
    a.ub2[o] = b2;
    a.ub3[o] = b3;
  } else {
    a.ub1[o] += b1;
    a.ub2[o] += b2;
    a.ub3[o] += b3;
  }
}

// k_adj_prod marching up the planes: one thread per (x, z) point and a chunk
// of J planes, with the j-1, j, j+1 values of every field the stencil reads
// held in a register window.  Each value then crosses HBM once per chunk
// (plus the chunk's two edge planes); with one thread per point the
// neighbour planes mostly missed L2, and the pass fetched 16.2 GB a launch
// against 8.7 GB of inputs at 512x257x512.  NT is the scalar count the
// windows are compiled for (the launcher takes it when nth == NT); the sums
// are those of k_adj_prod, term by term.
template <int NT>
__global__ void __launch_bounds__(256) k_adj_prod_march(AdjProdArgs a, Geo g, int J) {
  const long long pp = g.pp;
  const int n = g.ny2;
  const long long xz = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (xz >= pp) return;
  const int j0 = blockIdx.y * J, j1 = min(j0 + J, n);
  struct W {
    double m, c, p;  // planes j-1, j, j+1 (0 outside 0 .. n-1)
  };
  auto ld = [&](const double* f, int j) -> double { return (j >= 0 && j < n) ? f[(long long)j * pp + xz] : 0.0; };
  auto win = [&](const double* f) -> W { return W{ld(f, j0 - 1), ld(f, j0), ld(f, j0 + 1)}; };
  auto next = [&](W& w, const double* f, int j) {  // from plane j to j + 1
    w.m = w.c;
    w.c = w.p;
    w.p = ld(f, j + 2);
  };
  W U1 = win(a.u1), U2 = win(a.u2), U3 = win(a.u3);
  W B0 = win(a.B[0]), B1 = win(a.B[1]), B3 = win(a.B[3]);
  W B2 = win(a.B[2]), A3 = win(a.A[3]);  // planes j, j+1 only
  W TH[NT > 0 ? NT : 1], B4[NT > 0 ? NT : 1];
#pragma unroll
  for (int s = 0; s < NT; ++s) {
    TH[s] = win(a.th[s]);
    B4[s] = win(a.B[4 + 2 * s]);
  }
  for (int j = j0; j < j1; ++j) {
    const long long o = (long long)j * pp + xz;
    const bool hu = j + 1 < n, hd = j >= 1;
    const double u1c = U1.c, u2c = U2.c, u3c = U3.c;
    const double u1p = U1.p, u2p = U2.p, u3p = U3.p;
    const double u1m = U1.m, u2m = U2.m, u3m = U3.m;
    auto c_of = [&](int jj) { return 1.0 / (2.0 * a.dyf[jj]); };
    const double cj = c_of(j), cm = hd ? c_of(j - 1) : 0.0, cp = hu ? c_of(j + 1) : 0.0;
    const double A0 = a.A[0][o], A1 = a.A[1][o], A2 = a.A[2][o];
    double b1 = 0.0, b2 = 0.0, b3 = 0.0;
    // pair 0
    b1 += 2.0 * u1c * A0;
    b1 += cj * (u2p - u2c) * B0.c + cm * u2c * B0.m - cp * u2p * B0.p;
    b2 += -cj * (u1c + u1m) * B0.c + cm * (u1c + u1m) * B0.m;
    // pair 1
    b1 += u3c * A1;
    b3 += u1c * A1;
    b3 += cj * (u2p - u2c) * B1.c + cm * u2c * B1.m - cp * u2p * B1.p;
    b2 += -cj * (u3c + u3m) * B1.c + cm * (u3c + u3m) * B1.m;
    // pair 2
    b3 += 2.0 * u3c * A2;
    {
      const double i3 = hd ? (a.dyf[j] * u3c + a.dyf[j - 1] * u3m) / (2.0 * a.dy[j]) : 0.0;
      b2 += i3 * B2.c;
      b3 += (hd ? a.dyf[j] / (2.0 * a.dy[j]) : 0.0) * u2c * B2.c;
      if (hu) b3 += (a.dyf[j] / (2.0 * a.dy[j + 1])) * u2p * B2.p;
    }
    // pair 3
    {
      const double i1 = hd ? (a.dyf[j] * u1c + a.dyf[j - 1] * u1m) / (2.0 * a.dy[j]) : 0.0;
      b2 += i1 * A3.c;
      b1 += (hd ? a.dyf[j] / (2.0 * a.dy[j]) : 0.0) * u2c * A3.c;
      if (hu) b1 += (a.dyf[j] / (2.0 * a.dy[j + 1])) * u2p * A3.p;
      if (hd) b2 += (0.5 * (u2p - u2m) / a.dy[j]) * B3.c;
      if (hd) b2 += (0.5 * (u2m + u2c) / a.dy[j - 1 > 0 ? j - 1 : 1]) * B3.m;
      if (hu) b2 += (-0.5 * (u2p + u2c) / a.dy[j + 1]) * B3.p;
    }
    // scalars
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      const double tc = TH[s].c, tm = TH[s].m;
      const double A4 = a.A[4 + 2 * s][o], A5 = a.A[5 + 2 * s][o];
      double bt = 0.0;
      b1 += tc * A4;
      bt += u1c * A4;
      bt += cj * (u2p - u2c) * B4[s].c + cm * u2c * B4[s].m - cp * u2p * B4[s].p;
      b2 += -cj * (tc + tm) * B4[s].c + cm * (tc + tm) * B4[s].m;
      b3 += tc * A5;
      bt += u3c * A5;
      a.ubth[s][o] = a.store ? bt : a.ubth[s][o] + bt;
    }
    if (a.store) {
      a.ub1[o] = b1;
      a.ub2[o] = b2;
      a.ub3[o] = b3;
    } else {
      a.ub1[o] += b1;
      a.ub2[o] += b2;
      a.ub3[o] += b3;
    }
    next(U1, a.u1, j);
    next(U2, a.u2, j);
    next(U3, a.u3, j);
    next(B0, a.B[0], j);
    next(B1, a.B[1], j);
    next(B3, a.B[3], j);
    next(B2, a.B[2], j);
    next(A3, a.A[3], j);
#pragma unroll
    for (int s = 0; s < NT; ++s) {
      next(TH[s], a.th[s], j);
      next(B4[s], a.B[4 + 2 * s], j);
    }
  }
}

// The (0,0)-mode entries of a spectral field: a[j] (real part) over planes.
__global__ void k_add_mean(double2* __restrict__ c, const double* __restrict__ w, Geo g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.ny2) return;
  c[(long long)j * g.sp].x += w[j];
}

}  // namespace dgpu
