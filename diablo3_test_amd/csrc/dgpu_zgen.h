// z transforms of any length (Nz that is not a power of two, e.g. the
// 65-point z direction of the reference's front setups).
//
// The power-of-two passes hold a line in registers, 8 values per thread,
// and run radix-8 Stockham steps through LDS (dgpu_fft.h).  Here Nz is a
// product of arbitrary radices p_0 p_1 ... (small primes first), and each
// line is transformed in LDS by a mixed-radix Stockham recursion: pass i
// (radix p, Ns = p_0...p_{i-1}) maps the buffer a to b by
//   b[(j - k) p + k + s Ns] = sum_r a[j + r N/p] w_N^{r (N/(Ns p)) (k + s Ns)},
//   j = 0..N/p-1, k = j mod Ns, s = 0..p-1,
// i.e. the twiddle of the decimation-in-time step and the p-point DFT in one
// sum (O(p) per output, O(N sum p) per line).  w_N^e comes from the same
// table exp(-2 pi i e / N) as the power-of-two passes (conjugated for the
// backward transform).  The kernels mirror k_zfwd, k_zinv and k_zinv_op
// (same arguments, packing, scales and derivative ops); the spectral RHS
// pass reads the products' z spectra from k_zgen_fwd's output instead of
// transforming them itself (k_rhs with PRE).
#pragma once

#include "dgpu_kernels.h"
#include "dgpu_les.h"

namespace dgpu {

constexpr int ZGEN_MAXPASS = 24;

// Radices of Nz in pass order (factors 4 first, then 2, 3, 5, 7, ...).
struct ZPlan {
  int np;
  int rad[ZGEN_MAXPASS];
};

// Mixed-radix Stockham transform of the line in a (n values) by the nt
// threads t = 0..nt-1 of a group; b is scratch of n values.  Every thread of
// the block calls it with the same plan (it synchronises the block between
// passes); returns the buffer that holds the result.
template <typename R, int DIR>
__device__ cpx<R>* fft_gen(cpx<R>* a, cpx<R>* b, int n, const ZPlan& pl, const cpx<R>* __restrict__ W, int t,
                           int nt) {
  int Ns = 1;
  for (int ps = 0; ps < pl.np; ++ps) {
    const int p = pl.rad[ps], m = n / p, u = n / (Ns * p);
    for (int idx = t; idx < n; idx += nt) {
      const int j = idx % m, s = idx / m;
      const int k = j % Ns;
      const int step = u * (k + s * Ns);  // < n
      int e = 0;
      cpx<R> acc = {(R)0, (R)0};
      for (int r = 0; r < p; ++r) {
        cpx<R> w = W[e];
        if (DIR > 0) w.y = -w.y;
        const cpx<R> x = a[j + r * m];
        acc.x += x.x * w.x - x.y * w.y;
        acc.y += x.x * w.y + x.y * w.x;
        e += step;
        if (e >= n) e -= n;
      }
      b[(j - k) * p + k + s * Ns] = acc;
    }
    __syncthreads();
    cpx<R>* c = a;
    a = b;
    b = c;
    Ns *= p;
  }
  return a;
}

// Forward z transform + 1/Nz + pack, as k_zfwd.  Groups of tc threads, one
// column (j, q) each; LDS 2 Nz values per group.
template <typename R>
__global__ void __launch_bounds__(256) k_zgen_fwd(const cpx<R>* __restrict__ T, cpx<double>* __restrict__ C, Geo g,
                                                  ZPlan pl, const cpx<R>* __restrict__ W, int j0, int accum, R s0,
                                                  R s1, int tc, FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    T = reinterpret_cast<const cpx<R>*>(fb.in[blockIdx.z]);
    C = reinterpret_cast<cpx<double>*>(fb.out[blockIdx.z]);
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = g.nz;
  const int grp = threadIdx.x / tc, t = threadIdx.x % tc;
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw) + (long long)grp * 2 * n;
  const int j = j0 + blockIdx.y, q = blockIdx.x * (blockDim.x / tc) + grp;
  const bool act = q < g.nq;
  const cpx<R>* col = T + (long long)j * g.tp + (long long)(act ? q : 0) * n;
  for (int k = t; k < n; k += tc) a[k] = act ? ld_in(col + k) : cpx<R>{(R)0, (R)0};
  __syncthreads();
  const cpx<R>* res = fft_gen<R, -1>(a, a + n, n, pl, W, t, tc);
  if (act) {
    const R inv = ((q == 0) ? s0 : s1) / (R)n;
    cpx<double>* out = C + (long long)j * g.sp + (long long)q * g.nkk;
    for (int kk = t; kk < g.nkk; kk += tc) {
      const cpx<R> z = inv * res[kz_nat(g, kk)];
      cpx<double> zd = {(double)z.x, (double)z.y};
      if (accum) zd = out[kk] + zd;
      out[kk] = zd;
    }
  }
}

// Unpack + zero + backward z transform of a spectral field or of one of the
// derivative forms of k_zinv_op (op; OP_ID with the scales s0, s1 is
// k_zinv), unnormalised.
template <typename R>
__global__ void __launch_bounds__(256) k_zgen_inv(const cpx<double>* __restrict__ C, cpx<R>* __restrict__ T, Geo g,
                                                  ZPlan pl, const cpx<R>* __restrict__ W, int j0, int op,
                                                  const double* __restrict__ kx, const double* __restrict__ kz,
                                                  const double* __restrict__ dy, const double* __restrict__ dyf,
                                                  double s0, double s1, int tc, FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    C = reinterpret_cast<const cpx<double>*>(fb.in[blockIdx.z]);
    T = reinterpret_cast<cpx<R>*>(fb.out[blockIdx.z]);
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = g.nz;
  const int grp = threadIdx.x / tc, t = threadIdx.x % tc;
  cpx<R>* a = reinterpret_cast<cpx<R>*>(smem_raw) + (long long)grp * 2 * n;
  const int j = j0 + blockIdx.y, q = blockIdx.x * (blockDim.x / tc) + grp;
  bool act = q < g.nq;
  if ((op == OP_DY_GY || op == OP_DY_C2) && j == 0) act = false;
  if (op >= OP_DY_GYF && j == g.ny + 1) act = false;
  const cpx<double>* col = C + (long long)j * g.sp + (long long)(q < g.nq ? q : 0) * g.nkk;
  const double kxq = kx ? kx[q < g.nq ? q : 0] : 0.0;
  double sdy = 0.0;
  if (act && op == OP_DY_GY) sdy = 1.0 / dy[j];
  if (act && op == OP_DY_GYF) sdy = 1.0 / dyf[j];
  if (act && op == OP_DY_C2) sdy = 1.0 / (2.0 * dyf[j]);
  const double sc = (q == 0) ? s0 : s1;
  const int off = n - 2 * g.nkz - 1;
  for (int k = t; k < n; k += tc) {
    int kk = -1;
    if (k <= g.nkz)
      kk = k;
    else if (k >= n - g.nkz)
      kk = k - off;
    cpx<double> z = {0.0, 0.0};
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
    a[k] = {(R)(sc * z.x), (R)(sc * z.y)};
  }
  __syncthreads();
  const cpx<R>* res = fft_gen<R, 1>(a, a + n, n, pl, W, t, tc);
  if (q < g.nq) {
    cpx<R>* out = T + (long long)j * g.tp + (long long)q * n;
    for (int k = t; k < n; k += tc) st_out(out + k, res[k]);
  }
}

}  // namespace dgpu
