// dgpu_kernels.h -- HIP kernels of the channel RK3/Crank-Nicolson substep
// (reference: for/channel.f90 rk_chan_1, rem_div_chan; for/fft.f90).
//
// Device layouts (one GPU holds the whole field, so no pencils):
//   physical  P[j][k][i]       j = 0..Ny+1, k = 0..Nz-1, i = 0..Nx-1   (i fastest)
//   spectral  C[j][q][kk]      q = 0..Nkx (kx = q*2pi/Lx), kk = 0..2Nkz packed as
//                              the reference packs kz (fft.f90:307-312)   (kk fastest)
//   x-spectra T[j][q][k]       the half-way point of a 2-D transform: x already
//                              transformed and dealiased, z not yet     (k fastest)
// Only the dealiased kx block q <= Nkx is stored: the reference zeroes
// i in (Nkx, Nx/2] on every transform (fft.f90:264-266, 374-376) and no
// spectral operation of the substep couples kx modes, so those modes are
// identically zero in every field.
#pragma once
#include "dgpu_fft.h"

namespace dgpu {

// Rows (z lines) per workgroup in the x passes: each workgroup reads and
// writes its x-spectra as runs of KT complex values, one 128-byte line.
template <typename R>
struct XTile {
  static constexpr int KT = 128 / (2 * (int)sizeof(R));
};

// Whether the inverse x pass keeps its twiddle table in LDS (see k_xinv).
template <typename R>
struct XInvTable {
  static constexpr bool on = sizeof(R) == 4;
};

struct Geo {
  int nx, ny, nz;     // Nx, Ny, Nz
  int lognx, lognz;   // log2 Nx, log2 Nz
  int nkx, nkz;       // Nx/3, Nz/3 (domain.f90:18-19)
  int nq, nkk;        // Nkx+1, 2Nkz+1
  int ny2;            // Ny+2 planes
  long long pp;       // physical plane: Nx*Nz
  long long sp;       // spectral plane: nq*nkk
  long long tp;       // x-spectra plane: nq*Nz
};

// Plane of flat index idx in arrays of `plane` entries per plane.  Every
// field has fewer than 2^32 entries (checked when a handle is created), so a
// 32-bit division serves: a 64-bit one by a run-time divisor costs a
// per-thread subroutine of dozens of instructions.
__device__ __forceinline__ int plane_of(long long idx, long long plane) {
  return (int)((unsigned)idx / (unsigned)plane);
}

// natural z index of packed mode kk (inverse of the pack at fft.f90:308-311)
__device__ __forceinline__ int kz_nat(const Geo& g, int kk) {
  return kk <= g.nkz ? kk : g.nz - 2 * g.nkz - 1 + kk;
}

// Several fields through one transform pass in one launch: blockIdx.z picks
// the field's input and output (nf = 0: the single field of the plain
// arguments).  One launch keeps the CUs busy through each field's last blocks.
constexpr int MAXBF = 8;
struct FieldBatch {
  const void* in[MAXBF];
  void* out[MAXBF];
  int nf;
};

// ---------------------------------------------------------------- x pass --
// Real-pair post-processing: z = FFT(a + i b) -> A = FFT(a), B = FFT(b) at q.
template <typename R>
__device__ __forceinline__ void split_pair(const cpx<R>* f, int n, int q, R scale, cpx<R>& A,
                                           cpx<R>& B) {
  cpx<R> z = f[pidx(q)], zc = conjg(f[pidx((n - q) & (n - 1))]);
  R h = (R)0.5 * scale;
  A = {h * (z.x + zc.x), h * (z.y + zc.y)};
  B = {h * (z.y - zc.y), -h * (z.x - zc.x)};
}

// Forward x transform of one field on planes j0.. : the r2c of fft.f90:258
// with the 1/Nx scale and dealias of :259-268.  Two rows share one complex
// FFT; a workgroup owns KT rows of one plane and writes their kx <= Nkx
// spectra straight from the transform buffers as KT-long runs of T[j][q][k].
// grid (Nz/KT, nplanes), block ngroups*Nx/8 (ngroups divides KT/2),
// LDS ngroups*fft_lds_len(Nx) cpx.
// RAG: Nz is not a multiple of KT, so the last tile holds fewer rows; rows
// k >= Nz are read as zero and not stored (here and in the other x passes).
template <typename R, bool WAVE, bool RAG = false, int LOGN = 0>
__global__ void __launch_bounds__(512) k_xfwd(const R* __restrict__ u, cpx<R>* __restrict__ T,
                                               Geo g, const cpx<R>* __restrict__ tw, int j0,
                                               FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    u = reinterpret_cast<const R*>(fb.in[blockIdx.z]);
    T = reinterpret_cast<cpx<R>*>(fb.out[blockIdx.z]);
  }
  constexpr int KT = XTile<R>::KT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nx >> 3;
  const int L = x_group_stride(g.nx);
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * L;
  // the power-formed twiddles read one table entry a step: from the caches,
  // which leaves the table's LDS to the workgroups (4 instead of 3 a CU; the
  // LES x forward passes 6.52 -> 6.30 ms per step at 512x257x512)
  const cpx<R>* twl = LOGN == FFT_XL512P ? tw : lds_twiddles(smem + ngroups * L, tw, g.nx);
  const int j = j0 + blockIdx.y, k0 = blockIdx.x * KT;
  const R* plane = u + (long long)j * g.pp;
  cpx<R>* out = T + (long long)j * g.tp + k0;
  const R inv = (R)1 / (R)g.nx;
  const int nr = 2 * ngroups;  // rows per round
  for (int m0 = 0; m0 < KT / 2; m0 += ngroups) {
    const int m = m0 + grp;
    const R* ra = plane + (long long)(k0 + 2 * m) * g.nx;
    const R* rb = ra + g.nx;
    const bool oka = !RAG || k0 + 2 * m < g.nz, okb = !RAG || k0 + 2 * m + 1 < g.nz;
    cpx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      v[r] = {oka ? ld_in(ra + gt + G * r) : (R)0, okb ? ld_in(rb + gt + G * r) : (R)0};
    fft_any<R, -1, WAVE, LOGN>(v, fbuf, g.nx, g.lognx, twl, gt);
    __syncthreads();  // every group's spectra are read below
    for (int idx = threadIdx.x; idx < g.nq * nr; idx += blockDim.x) {
      const int q = idx / nr, kk = idx - q * nr;
      if (RAG && k0 + 2 * m0 + kk >= g.nz) continue;
      cpx<R> A, B;
      split_pair(smem + (kk >> 1) * L, g.nx, q, inv, A, B);
      st_out(out + (long long)q * g.nz + 2 * m0 + kk, (kk & 1) ? B : A);
    }
    __syncthreads();
  }
}

// Inverse x transform: dealias (fft.f90:372-378) and c2r (:380), unnormalised.
// FFTW's c2r reads only the real part of the kx = 0 input.  Rows 2m and 2m+1
// are synthesised together: Z = A + iB with A, B extended Hermitian.  With dkx
// the x-spectra are multiplied by i kx first (the x derivative).
template <typename R, bool WAVE, int LOGN = 0, bool RAG = false>
__global__ void __launch_bounds__(512) k_xinv(const cpx<R>* __restrict__ T, R* __restrict__ u,
                                               Geo g, const cpx<R>* __restrict__ tw, int j0,
                                               const double* __restrict__ dkx = nullptr,
                                               FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    T = reinterpret_cast<const cpx<R>*>(fb.in[blockIdx.z]);
    u = reinterpret_cast<R*>(fb.out[blockIdx.z]);
  }
  constexpr int KT = XTile<R>::KT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nx >> 3;
  const int L = x_group_stride(g.nx);
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * L;
  // twiddles in LDS for the FP32 pass only: in FP64 the 8 KB table costs a
  // workgroup per CU (4 -> 3 at 512 points), and the pass went from 2.32 to
  // 2.50 ms per step at 512x257x512 with it, against 1.39 -> 1.11 in FP32
  const cpx<R>* twl = XInvTable<R>::on ? lds_twiddles(smem + ngroups * L, tw, g.nx) : tw;
  const int j = j0 + blockIdx.y, k0 = blockIdx.x * KT;
  const cpx<R>* in = T + (long long)j * g.tp + k0;
  R* plane = u + (long long)j * g.pp;
  for (int m0 = 0; m0 < KT / 2; m0 += ngroups) {
    __syncthreads();  // previous round's reads of the buffers are done
    for (int qq = g.nkx + 1 + gt; qq < g.nx - g.nkx; qq += G) fbuf[pidx(qq)] = {(R)0, (R)0};
    for (int idx = threadIdx.x; idx < g.nq * ngroups; idx += blockDim.x) {
      const int q = idx / ngroups, mm = idx - q * ngroups;
      const cpx<R>* src = in + (long long)q * g.nz + 2 * (m0 + mm);
      // streaming loads pay off for FP32 spectra (3.46 -> 2.90 ms per step at
      // 512x257x512) but not for FP64 ones (4.89 -> 5.40 ms)
      cpx<R> A, B;
      if (RAG) {
        const int ka = k0 + 2 * (m0 + mm);
        A = ka < g.nz ? src[0] : cpx<R>{(R)0, (R)0};
        B = ka + 1 < g.nz ? src[1] : cpx<R>{(R)0, (R)0};
      } else if (sizeof(R) == 4) {
        A = ld_in(src);
        B = ld_in(src + 1);
      } else {
        A = src[0];
        B = src[1];
      }
      if (dkx) {
        const R k = (R)dkx[q];
        A = ik(k, A);
        B = ik(k, B);
      }
      cpx<R>* f = smem + mm * L;
      if (q == 0) {
        f[pidx(0)] = {A.x, B.x};
      } else {
        f[pidx(q)] = {A.x - B.y, A.y + B.x};
        f[pidx(g.nx - q)] = {A.x + B.y, B.x - A.y};
      }
    }
    __syncthreads();  // the spectra of every group are in place
    cpx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = fbuf[pidx(gt + G * r)];
    fft_any<R, 1, WAVE, LOGN>(v, fbuf, g.nx, g.lognx, twl, gt);
    R* ra = plane + (long long)(k0 + 2 * (m0 + grp)) * g.nx;
    R* rb = ra + g.nx;
    const bool oka = !RAG || k0 + 2 * (m0 + grp) < g.nz, okb = !RAG || k0 + 2 * (m0 + grp) + 1 < g.nz;
    for (int i = gt; i < g.nx; i += G) {
      const cpx<R> z = fbuf[pidx(i)];
      if (oka) st_out(ra + i, z.x);
      if (okb) st_out(rb + i, z.y);
    }
  }
}

// ---------------------------------------------------------------- z pass --
// Forward z transform + 1/Nz + pack (fft.f90:298-313).  One column (j, q) per
// group; grid (ceil(nq/ngroups), nplanes), block ngroups*Nz/8.  The spectral
// array is FP64 whatever the transform precision R; with accum the result is
// added to it (the FP32 transform of an increment of an FP64 state).
template <typename R, bool WAVE, int LOGN = 0>
__global__ void __launch_bounds__(512) k_zfwd(const cpx<R>* __restrict__ T, cpx<double>* __restrict__ C,
                                               Geo g, const cpx<R>* __restrict__ tw, int j0, int accum,
                                               R s0, R s1, FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    T = reinterpret_cast<const cpx<R>*>(fb.in[blockIdx.z]);
    C = reinterpret_cast<cpx<double>*>(fb.out[blockIdx.z]);
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nz >> 3;
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * fft_lds_len(g.nz);
  // power-formed twiddles (one table entry a step) from the caches, the
  // table's LDS left to the workgroups
  const cpx<R>* twl = LOGN == FFT_XL512P ? tw : lds_twiddles(smem + ngroups * fft_lds_len(g.nz), tw, g.nz);
  const int j = j0 + blockIdx.y, q = blockIdx.x * ngroups + grp;
  const bool act = q < g.nq;
  cpx<R> v[8];
  if (act) {
    const cpx<R>* col = T + (long long)j * g.tp + (long long)q * g.nz;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = ld_in(col + gt + G * r);
  } else {
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = {(R)0, (R)0};
  }
  fft_any<R, -1, WAVE, LOGN>(v, fbuf, g.nz, g.lognz, twl, gt);
  if (act) {
    // 1/Nz, times s0 (kx = 0) or s1 (kx > 0): 1 for the transform itself,
    // Nx Nz W' for the transpose of the inverse transform
    const R inv = ((q == 0) ? s0 : s1) / (R)g.nz;
    cpx<double>* out = C + (long long)j * g.sp + (long long)q * g.nkk;
    for (int kk = gt; kk < g.nkk; kk += G) {
      const cpx<R> z = inv * fbuf[pidx(kz_nat(g, kk))];
      cpx<double> zd = {(double)z.x, (double)z.y};
      if (accum) zd = out[kk] + zd;
      out[kk] = zd;  // plain store: the projection and the next substep read the state
    }
  }
}

// Unpack + zero (fft.f90:330-341) + backward z transform (:345), unnormalised.
// s0 (kx = 0) and s1 (kx > 0) scale the input: 1 for the transform, W/(Nx Nz)
// for the transpose of the forward transform.
template <typename R, bool WAVE, int LOGN = 0>
__global__ void __launch_bounds__(512) k_zinv(const cpx<double>* __restrict__ C, cpx<R>* __restrict__ T,
                                               Geo g, const cpx<R>* __restrict__ tw, int j0, double s0,
                                               double s1, FieldBatch fb = FieldBatch{}) {
  if (fb.nf > 0) {
    C = reinterpret_cast<const cpx<double>*>(fb.in[blockIdx.z]);
    T = reinterpret_cast<cpx<R>*>(fb.out[blockIdx.z]);
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nz >> 3;
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * fft_lds_len(g.nz);
  // power-formed twiddles (one table entry a step) from the caches, the
  // table's LDS left to the workgroups
  const cpx<R>* twl = LOGN == FFT_XL512P ? tw : lds_twiddles(smem + ngroups * fft_lds_len(g.nz), tw, g.nz);
  const int j = j0 + blockIdx.y, q = blockIdx.x * ngroups + grp;
  const bool act = q < g.nq;
  cpx<R> v[8];
  const cpx<double>* col = C + (long long)j * g.sp + (long long)(act ? q : 0) * g.nkk;
  const int off = g.nz - 2 * g.nkz - 1;
  const double sc = (q == 0) ? s0 : s1;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int k = gt + G * r;
    cpx<double> z = {0.0, 0.0};
    if (act) {
      if (k <= g.nkz)
        z = ld_in(col + k);
      else if (k >= g.nz - g.nkz)
        z = ld_in(col + k - off);
    }
    v[r] = {(R)(sc * z.x), (R)(sc * z.y)};
  }
  fft_any<R, 1, WAVE, LOGN>(v, fbuf, g.nz, g.lognz, twl, gt);
  if (act) {
    cpx<R>* out = T + (long long)j * g.tp + (long long)q * g.nz;
    for (int k = gt; k < g.nz; k += G) st_out(out + k, fbuf[pidx(k)]);
  }
}

// ------------------------------------------------------ nonlinear x pass --
// The products of channel.f90:299-459 (velocity) and :557-608 (scalar n),
// computed on the fly from the physical fields and x-transformed in pairs.
// Each pair feeds one or two x-spectra; terms that the reference multiplies
// by ikx after its 2-D transform get that factor here (kx commutes with the z
// transform), terms multiplied by ikz keep a spectrum of their own:
//   pair 0: (u1u1, V1)   -> X1  = ikx F(u1u1) + F(V1)            -> cf1
//   pair 1: (u3u1, V3)   -> X13 = F(u3u1); X3 = ikx F(u3u1) + F(V3)
//   pair 2: (u3u3, W3)   -> X33 = F(u3u3); X32 = F(W3)            (W3 = interp(u3) u2)
//   pair 3: (W1, V2)     -> X2  = ikx F(W1) + F(V2)              (W1 = interp(u1) u2)
//   pair 4: (th u1, Vt)  -> Xt  = ikx F(th u1) + F(Vt)
//   pair 5: (th u3, 0)   -> Xt3 = F(th u3)
constexpr int MAXTH = 16;  // DGPU_MAX_TH; k_rhs keeps a bit per scalar in a 32-bit mask
static_assert(MAXTH <= 32, "k_rhs holds the scalars' row flags in one 32-bit mask");
constexpr int MAXSLOT = 6 + 2 * MAXTH;

struct ProdArgs {
  const void* u1;    // physical fields in the transform precision R
  const void* u2;
  const void* u3;
  const double* dy;
  const double* dyf;
  const double* kx;  // kx[q]
  int js, je;        // jstart, jend
  int jst, jet;      // jstart_th(n), jend_th(n)
};

// All product pairs of one substep in one pass (the slot numbering of k_rhs):
// pair 0 -> slot 0, pair 1 -> slots 1, 2, pair 2 -> slots 3, 5, pair 3 -> slot 4,
// scalar n: pair 4 -> slot 6+2n, pair 5 -> slot 7+2n.
struct ProdAllArgs {
  ProdArgs base;
  const void* th[MAXTH];
  int jst[MAXTH], jet[MAXTH];
  int nth;
  double2* out[MAXSLOT];
};

__device__ __forceinline__ void pair_ranges(const ProdArgs& a, const Geo& g, int pair, int j,
                                            bool& inA, bool& inB) {
  const bool r1 = j >= a.js && j <= a.je;
  const bool r2 = j >= 2 && j <= g.ny;
  const bool rt = j >= a.jst && j <= a.jet;
  switch (pair) {
    case 0:
    case 1: inA = r1; inB = r1; break;
    case 2: inA = r1; inB = r2; break;
    case 3: inA = r2; inB = r2; break;
    case 4: inA = rt; inB = rt; break;
    default: inA = rt; inB = false; break;
  }
}

// All product pairs for a KT-row tile of plane j in one workgroup.  Each
// thread reads its slice of the input rows (planes j-1, j, j+1) once into
// registers, in three phases -- u1 u2 (pairs 0, 3), u3 (pairs 1, 2), then
// each scalar (pairs 4, 5) -- and every product is formed from registers, so
// each input row crosses the memory system once per workgroup.
template <typename R>
__device__ __forceinline__ void emit_pair(cpx<R>* smem, int L, const Geo& g, int lg_ng, int pair,
                                          R inv, const R* kx, cpx<R>* out0, cpx<R>* out1,
                                          int m0, int nrow = 1 << 30) {
  const int ngroups = 1 << lg_ng;
  for (int idx = threadIdx.x; idx < g.nq * ngroups; idx += blockDim.x) {
    const int q = idx >> lg_ng, mm = idx & (ngroups - 1);
    if (mm >= nrow) continue;  // rows past Nz of a ragged tile
    cpx<R> A, B;
    split_pair(smem + mm * L, g.nx, q, inv, A, B);
    const cpx<R> ikA = ik(kx[q], A);
    const int o = q * g.nz + m0 + mm;  // within one plane's x-spectra
    switch (pair) {
      case 0: st_out(out0 + o, ikA + B); break;
      case 1: st_out(out0 + o, A); st_out(out1 + o, ikA + B); break;
      case 2: st_out(out0 + o, A); st_out(out1 + o, B); break;
      case 3:
      case 4: st_out(out0 + o, ikA + B); break;
      default: st_out(out0 + o, A); break;
    }
  }
}

// LOGNX: Nx at compile time as in fft_any (fft_regs_c, or fft512_xl for
// FFT_XL512); 0: Nx from g.
// PF: each input phase's loads are issued into the registers the previous
// pair has just finished with, before that pair's transform, so that they
// are in flight while it runs (the inputs still cross the memory system once).
template <typename R, bool WAVE, int LOGNX = 0, bool RAG = false, bool PF = false>
__global__ void __launch_bounds__(512) k_xprod_all(ProdAllArgs a, Geo g,
                                                    const cpx<R>* __restrict__ tw, int j0) {
  constexpr int KT = XTile<R>::KT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int G = g.nx >> 3;  // 8 values of a row per thread
  const int L = x_group_stride(g.nx);
  const int ngroups = blockDim.x / G;  // a power of two
  const int lg_ng = 31 - __builtin_clz(ngroups);
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  cpx<R>* fbuf = smem + grp * L;
  // the table stays in LDS also for the power-formed twiddles (one entry a
  // step): read from the caches instead, the FP64 pass went 4.70 -> 5.18 ms
  // per step at 512x257x512 (its registers, not LDS, bound its occupancy)
  const ProdArgs& b = a.base;
  // kx in LDS behind the table: the emit loops then hold no global load, whose
  // wait would also wait for the stores of the previous items
  R* kxl = reinterpret_cast<R*>(smem + ngroups * L + g.nx);
  for (int t = threadIdx.x; t < g.nq; t += blockDim.x) kxl[t] = (R)b.kx[t];
  const cpx<R>* twl = lds_twiddles(smem + ngroups * L, tw, g.nx);  // synchronises
  const int j = j0 + blockIdx.y, k0 = blockIdx.x * KT;
  const R inv = (R)1 / (R)g.nx;
  // plane j of each field (uniform across the block); per-thread offsets
  // within a plane are 32-bit
  const long long pp = g.pp;
  const R* U1 = reinterpret_cast<const R*>(b.u1);
  const R* U2 = reinterpret_cast<const R*>(b.u2);
  const R* U3 = reinterpret_cast<const R*>(b.u3);
  const R dyj = (R)b.dy[j], dyfj = (R)b.dyf[j], dyfm = (R)b.dyf[j - 1];
  const R c2dyf = (R)1 / ((R)2 * dyfj);
  const R c2dy = (R)1 / ((R)2 * dyj), cdy = (R)1 / dyj;
  const R quarter = (R)0.25, zero = (R)0;
  auto outp = [&](int s) {
    return reinterpret_cast<cpx<R>*>(a.out[s]) + (long long)j * g.tp + k0;
  };
  // transform one product pair held as this thread's 8 values, then make
  // every group's spectra visible to the cross-group writes of emit_pair
  auto fft_pair = [&](const R* va, const R* vb) {
    cpx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = {va[r], vb[r]};
    fft_any<R, -1, WAVE, LOGNX>(v, fbuf, g.nx, g.lognx, twl, gt);
    __syncthreads();
  };
  bool r1, r2;
  {
    bool x;
    pair_ranges(b, g, 0, j, r1, x);
    pair_ranges(b, g, 3, j, r2, x);
  }
  if constexpr (PF) {
    for (int m0 = 0; m0 < KT; m0 += ngroups) {
      const bool rowok = !RAG || k0 + m0 + grp < g.nz;
      const int nrow = RAG ? g.nz - k0 - m0 : 1 << 30;
      const long long rowc = (long long)j * pp + (long long)(rowok ? k0 + m0 + grp : 0) * g.nx + gt;
      auto ld8 = [&](const R* F, long long off, R(&d)[8]) {
#pragma unroll
        for (int r = 0; r < 8; ++r) d[r] = F[rowc + off + G * r];
      };
      auto thp_of = [&](int n) { return reinterpret_cast<const R*>(a.th[n]); };
      // register sets, by what they hold in turn:
      //   A: u1(j-1), u3(j-1), th(j-1)    B: u1(j)     Cs: u1(j+1), u3(j+1), th(j+1)
      //   D: u2(j-1), u3(j)               E: u2(j)     F: u2(j+1)       T: th(j)
      R A[8], B[8], Cs[8], D[8], E[8], F[8], T[8];
      R va[8], vb[8];
      ld8(U1, -pp, A);
      ld8(U1, 0, B);
      ld8(U1, pp, Cs);
      ld8(U2, -pp, D);
      ld8(U2, 0, E);
      ld8(U2, pp, F);
      // pair 0: (u1u1, V1) -> slot 0                                    (channel.f90:319-325, 399-407)
      if (r1) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          va[r] = B[r] * B[r];
          vb[r] = (Cs[r] * F[r] + B[r] * F[r] - B[r] * E[r] - A[r] * E[r]) * c2dyf;
        }
      }
      ld8(U3, pp, Cs);  // u1(j+1) is done with
      if (r1) {
        fft_pair(va, vb);
        emit_pair(smem, L, g, lg_ng, 0, inv, kxl, outp(0), (cpx<R>*)nullptr, m0, nrow);
        __syncthreads();
      }
      // pair 3: (W1, V2) -> slot 4                                      (:356-365, 440-449)
      if (r2) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          va[r] = ((dyfj * B[r] + dyfm * A[r]) * c2dy) * E[r];
          const R sp = E[r] + F[r], sm = E[r] + D[r];
          vb[r] = (quarter * (sp * sp) - quarter * (sm * sm)) * cdy;
        }
      }
      ld8(U3, -pp, A);  // u1(j-1) and u2(j-1) are done with
      ld8(U3, 0, D);
      if (r2) {
        fft_pair(va, vb);
        emit_pair(smem, L, g, lg_ng, 3, inv, kxl, outp(4), (cpx<R>*)nullptr, m0, nrow);
        __syncthreads();
      }
      // pair 1: (u3u1, V3) -> slots 1, 2                                (:298-305, 419-428)
      if (r1) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          va[r] = D[r] * B[r];
          vb[r] = (Cs[r] * F[r] + D[r] * F[r] - D[r] * E[r] - A[r] * E[r]) * c2dyf;
        }
      }
      if (a.nth > 0) {  // u3(j+1) is done with
        ld8(thp_of(0), pp, Cs);
        ld8(thp_of(0), 0, T);
      }
      if (r1) {
        fft_pair(va, vb);
        emit_pair(smem, L, g, lg_ng, 1, inv, kxl, outp(1), outp(2), m0, nrow);
        __syncthreads();
      }
      // pair 2: (u3u3, W3) -> slots 3, 5                                (:337-344, 377-386)
      if (r1 || r2) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          va[r] = r1 ? D[r] * D[r] : zero;
          vb[r] = r2 ? ((dyfj * D[r] + dyfm * A[r]) * c2dy) * E[r] : zero;
        }
      }
      if (a.nth > 0) ld8(thp_of(0), -pp, A);  // u3(j-1) is done with
      if (r1 || r2) {
        fft_pair(va, vb);
        emit_pair(smem, L, g, lg_ng, 2, inv, kxl, outp(3), outp(5), m0, nrow);
        __syncthreads();
      }
      // scalars: (th u1, Vt) -> slot 6+2n, (th u3, 0) -> slot 7+2n     (:557-600)
      // scalar n is in A, T, Cs; the next one's rows load behind its pairs
      for (int n = 0; n < a.nth; ++n) {
        const bool act = j >= a.jst[n] && j <= a.jet[n];  // uniform across the block
        const bool more = n + 1 < a.nth;
        if (act) {
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            va[r] = T[r] * B[r];
            vb[r] = (Cs[r] * F[r] + T[r] * F[r] - T[r] * E[r] - A[r] * E[r]) * c2dyf;
          }
        }
        if (more) {
          ld8(thp_of(n + 1), pp, Cs);
          ld8(thp_of(n + 1), -pp, A);
        }
        if (act) {
          fft_pair(va, vb);
          emit_pair(smem, L, g, lg_ng, 4, inv, kxl, outp(6 + 2 * n), (cpx<R>*)nullptr, m0, nrow);
          __syncthreads();
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            va[r] = T[r] * D[r];
            vb[r] = zero;
          }
        }
        if (more) ld8(thp_of(n + 1), 0, T);
        if (act) {
          fft_pair(va, vb);
          emit_pair(smem, L, g, lg_ng, 5, inv, kxl, outp(7 + 2 * n), (cpx<R>*)nullptr, m0, nrow);
          __syncthreads();
        }
      }
    }
    return;
  }
  for (int m0 = 0; m0 < KT; m0 += ngroups) {
    // a row past Nz (ragged tile) reads row 0 and is not stored
    const bool rowok = !RAG || k0 + m0 + grp < g.nz;
    const int nrow = RAG ? g.nz - k0 - m0 : 1 << 30;
    const long long rowc = (long long)j * pp + (long long)(rowok ? k0 + m0 + grp : 0) * g.nx + gt;
    R u1m[8], u1c[8], u1p[8], u2m[8], u2c[8], u2p[8];
    R va[8], vb[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const long long o = rowc + G * r;
      u1m[r] = U1[o - pp];
      u1c[r] = U1[o];
      u1p[r] = U1[o + pp];
      u2m[r] = U2[o - pp];
      u2c[r] = U2[o];
      u2p[r] = U2[o + pp];
    }
    // pair 0: (u1u1, V1) -> slot 0 = ikx F(u1u1) + F(V1)              (channel.f90:319-325, 399-407)
    if (r1) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = u1c[r] * u1c[r];
        vb[r] = (u1p[r] * u2p[r] + u1c[r] * u2p[r] - u1c[r] * u2c[r] - u1m[r] * u2c[r]) * c2dyf;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 0, inv, kxl, outp(0), (cpx<R>*)nullptr, m0, nrow);
      __syncthreads();
    }
    // pair 3: (W1, V2) -> slot 4 = ikx F(W1) + F(V2)                   (:356-365, 440-449)
    if (r2) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = ((dyfj * u1c[r] + dyfm * u1m[r]) * c2dy) * u2c[r];
        const R sp = u2c[r] + u2p[r], sm = u2c[r] + u2m[r];
        vb[r] = (quarter * (sp * sp) - quarter * (sm * sm)) * cdy;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 3, inv, kxl, outp(4), (cpx<R>*)nullptr, m0, nrow);
      __syncthreads();
    }
    // u3 at j-1, j, j+1 replaces u1 at j-1, j+1 and u2 at j-1
    R* u3m = u1m;
    R* u3p = u1p;
    R* u3c = u2m;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const long long o = rowc + G * r;
      u3m[r] = U3[o - pp];
      u3c[r] = U3[o];
      u3p[r] = U3[o + pp];
    }
    // pair 1: (u3u1, V3) -> slot 1 = F(u3u1), slot 2 = ikx F(u3u1) + F(V3)   (:298-305, 419-428)
    if (r1) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = u3c[r] * u1c[r];
        vb[r] = (u3p[r] * u2p[r] + u3c[r] * u2p[r] - u3c[r] * u2c[r] - u3m[r] * u2c[r]) * c2dyf;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 1, inv, kxl, outp(1), outp(2), m0, nrow);
      __syncthreads();
    }
    // pair 2: (u3u3, W3) -> slot 3 = F(u3u3), slot 5 = F(W3)            (:337-344, 377-386)
    if (r1 || r2) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = r1 ? u3c[r] * u3c[r] : zero;
        vb[r] = r2 ? ((dyfj * u3c[r] + dyfm * u3m[r]) * c2dy) * u2c[r] : zero;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 2, inv, kxl, outp(3), outp(5), m0, nrow);
      __syncthreads();
    }
    // scalars: (th u1, Vt) -> slot 6+2n, (th u3, 0) -> slot 7+2n       (:557-600)
    for (int n = 0; n < a.nth; ++n) {
      if (j < a.jst[n] || j > a.jet[n]) continue;  // uniform across the block
      const R* th = reinterpret_cast<const R*>(a.th[n]);
      R thm[8], thc[8], thp[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const long long o = rowc + G * r;
        thm[r] = th[o - pp];
        thc[r] = th[o];
        thp[r] = th[o + pp];
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = thc[r] * u1c[r];
        vb[r] = (thp[r] * u2p[r] + thc[r] * u2p[r] - thc[r] * u2c[r] - thm[r] * u2c[r]) * c2dyf;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 4, inv, kxl, outp(6 + 2 * n), (cpx<R>*)nullptr, m0, nrow);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        va[r] = thc[r] * u3c[r];
        vb[r] = zero;
      }
      fft_pair(va, vb);
      emit_pair(smem, L, g, lg_ng, 5, inv, kxl, outp(7 + 2 * n), (cpx<R>*)nullptr, m0, nrow);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------- spectral RHS z pass --
// For one (j, q) column: z-transform the x-spectra of the products, assemble
// the explicit RK terms of channel.f90:39-253 and :299-482 (velocity) and
// :209-246, :557-617 (scalars), store the new cf, form cr, and take cr back
// through the backward z transform (first half of fft_xz_to_physical, :485-487,
// :621).  Slots: 0 X1, 1 X13, 2 X3, 3 X33, 4 X2, 5 X32, 6+2n Xt_n, 7+2n Xt3_n.
// cr goes to slot 0 (cr1), 1 (cr3), 2 (cr2), 3+n (crth_n) once read.
struct RhsArgs {
  const void* X[MAXSLOT];         // x-spectra of the products (transform precision)
  void* Tcr[3 + MAXTH];           // out: x-spectra of cr1, cr3, cr2, crth_n (slot order)
  // out, instead of Tcr when set: cr itself (FP64 spectra, same slot order)
  // for the CN solve in spectral space; no backward z transform is run.
  // With crs32 (the mixed form) they hold cr less the state as FP32 spectra
  // (with incr = 2), which the solve adds back to the FP64 state.
  double2* crs[3 + MAXTH];
  int crs32;
  const double2* cu1;             // state the explicit terms read
  const double2* cu2;
  const double2* cu3;
  const double2* cp;
  const double2* cth[MAXTH];
  // With LES: planes 0, 1, Ny, Ny+1 of cu1..3, cth as they were before the
  // LES stage's wall values (4 planes each, pitch sp); the explicit terms read
  // those planes from here and the others from cu*, cth, which the LES stage
  // changes only on those planes.  Null otherwise.
  const double2* wcu1;
  const double2* wcu2;
  const double2* wcu3;
  const double2* wcth[MAXTH];
  // incr: cr is written less the state that was transformed to physical space
  // (cub*), so that only the O(h) increment passes through the transform
  int incr;  // 0: cr; 1: cr less cub*; 2: cr less the state read (cub* == cu*)
  const double2* cub1;
  const double2* cub2;
  const double2* cub3;
  const double2* cthb[MAXTH];
  // the RK carry, stored in the transform precision R (FP32 in the mixed form)
  double2* cf1;
  double2* cf2;
  double2* cf3;
  double2* cfth[MAXTH];
  const double* kx;
  const double* kz;
  const double* dy;
  const double* dyf;
  double nu, beta, coriolis;  // coriolis = Ro_inv/delta
  double grav_x, grav_y, grav_z;
  double pr[MAXTH], dthdx[MAXTH], dthdz[MAXTH];
  const double2* lesf1;             // subgrid forcing of cf1..3, cfth (null without LES)
  const double2* lesf2;
  const double2* lesf3;
  const double2* lesfth[MAXTH];
  double temp3, temp4, temp5, force;  // force = px0 (+ oscillation), 0 if none
  // thermal-wind advection of the Front flavour (forcing.f90:50-100), null
  // otherwise: fr_yf[j] = (gyf(j) - Ly/2) delta/Ro_inv, fr_y[j] the same at
  // gy(j); fr_sx, fr_sz the sums of dthdx, dthdz over the scalars (the
  // velocity terms of every scalar add up), fr_dr = delta/Ro_inv
  const double* fr_yf;
  const double* fr_y;
  double fr_sx, fr_sz, fr_dr;
  int rk_step, nth;
  int js, je;
  int jst[MAXTH], jet[MAXTH];
  // The outputs of this launch.  part & 1: cr1, cr3 (product slots 0-3);
  // part & 2: cr2 (slots 4, 5); scalars th0 <= n < th1 (slots 6+2n, 7+2n).
  // The LDS holds one buffer per slot of the launch, so a substep whose
  // slots do not all fit is split over launches with disjoint outputs.
  int part, th0, th1;
  int tw_lds;  // 1: the twiddle table is copied to LDS (0: read from HBM)
};

// PRE: the product slots X[s] already hold their z spectra (FP64, packed,
// scaled by 1/Nz: k_zgen_fwd for z lengths that are not powers of two); the
// pass then only combines them, and cr goes to crs (the caller transforms
// it further where needed).
template <typename R, bool WAVE, int LOGN = 0, bool PRE = false>
__global__ void __launch_bounds__(512) k_rhs(RhsArgs a, Geo g, const cpx<R>* __restrict__ tw,
                                              int j0) {
  typedef cpx<double> C;
  typedef cpx<R> CR;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  CR* smem = reinterpret_cast<CR*>(smem_raw);
  const int G = g.nz >> 3;
  const int ngroups = blockDim.x / G;
  const int grp = threadIdx.x / G, gt = threadIdx.x % G;
  const int j = j0 + blockIdx.y, q = blockIdx.x;
  const int L = fft_lds_len(g.nz);
  const bool spec = a.crs[0] != nullptr;
  const bool p13 = (a.part & 1) != 0, p2 = (a.part & 2) != 0;
  const bool in1 = p13 && j >= a.js && j <= a.je;
  const bool in2 = p2 && j >= 2 && j <= g.ny;
  unsigned inT = 0;  // bit n: scalar n of this launch has row j
  for (int n = a.th0; n < a.th1; ++n)
    if (j >= a.jst[n] && j <= a.jet[n]) inT |= 1u << n;
  // local slots (LDS buffers) of the launch: its velocity slots, then the
  // pairs of its scalars; cr slot c (0 cr1, 1 cr3, 2 cr2, 3+n crth_n) goes
  // to local buffer lcr(c), whose products its own outputs have read first
  const int nv = (p13 ? 4 : 0) + (p2 ? 2 : 0);
  const int nslot = nv + 2 * (a.th1 - a.th0);
  const int ncv = (p13 ? 2 : 0) + (p2 ? 1 : 0);
  const int ncr = ncv + (a.th1 - a.th0);
  auto gslot = [&](int ls) -> int { return ls < nv ? (p13 ? ls : 4 + ls) : 6 + 2 * a.th0 + (ls - nv); };
  auto lslot = [&](int s) -> int { return s < 4 ? s : s < 6 ? s - 4 + (p13 ? 4 : 0) : nv + s - 6 - 2 * a.th0; };
  auto gcr = [&](int lc) -> int { return lc < ncv ? (p13 ? lc : 2) : 3 + a.th0 + (lc - ncv); };
  auto lcr = [&](int c) -> int { return c < 2 ? c : c == 2 ? (p13 ? 2 : 0) : ncv + c - 3 - a.th0; };
  auto slot_used = [&](int s) -> bool {
    if (s < 4) return in1;
    if (s < 6) return in2;
    return (inT >> ((s - 6) >> 1)) & 1u;
  };
  const long long to = (long long)j * g.tp + (long long)q * g.nz;
  const int rounds = (nslot + ngroups - 1) / ngroups;
  const CR* twl = (!PRE && a.tw_lds) ? lds_twiddles(smem + (long long)rounds * ngroups * L, tw, g.nz) : tw;
  // --- z-transform the product spectra
  // The LDS holds one buffer per (round, group), so groups without a slot in
  // a round transform an unused buffer and every thread meets every barrier.
  for (int s0 = 0; s0 < (PRE ? 0 : nslot); s0 += ngroups) {
    const int ls = s0 + grp;
    const bool act = ls < nslot && slot_used(gslot(ls));
    CR* buf = smem + (long long)ls * L;
    CR v[8];
    if (act) {
      const CR* col = reinterpret_cast<const CR*>(a.X[gslot(ls)]) + to;
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = col[gt + G * r];
    } else {
#pragma unroll
      for (int r = 0; r < 8; ++r) v[r] = CR{(R)0, (R)0};
    }
    fft_any<R, -1, WAVE, LOGN>(v, buf, g.nz, g.lognz, twl, gt);
  }
  __syncthreads();  // the combine below reads every slot
  // --- explicit terms, mode by mode
  const double invnz = 1.0 / g.nz;
  const double kxq = a.kx[q];
  const double cdy = a.dy[j], cdyf = a.dyf[j], cdyfm = a.dyf[j > 0 ? j - 1 : 0];
  for (int kk = threadIdx.x; kk < g.nkk; kk += blockDim.x) {
    const int nat = kz_nat(g, kk);
    const long long o = (long long)j * g.sp + (long long)q * g.nkk + kk;
    const double kzk = a.kz[kk];
    const double lap = kxq * kxq + kzk * kzk;
    const double visc = (a.beta == 1.0) ? lap : pow(lap, a.beta);
    const bool mean = (q == 0 && kk == 0);
    // slot s at this mode; a slot is read before the cr slot that reuses it
    // is written (cr slot 3+n < 6+2n, and slots 0..3 belong to u1/u3)
    const int pn = pidx(nat);
    auto Z = [&](int s) -> C {
      if constexpr (PRE) {
        return reinterpret_cast<const C*>(a.X[s])[o];
      } else {
        const CR z = smem[(long long)lslot(s) * L + pn];
        return invnz * C{(double)z.x, (double)z.y};
      }
    };
    auto put = [&](int c, C v) {
      if (spec && a.crs32)
        reinterpret_cast<cpx<float>*>(a.crs[c])[o] = cpx<float>{(float)v.x, (float)v.y};
      else if (spec)
        reinterpret_cast<C*>(a.crs[c])[o] = v;
      else
        smem[(long long)lcr(c) * L + pn] = CR{(R)v.x, (R)v.y};
    };
    auto ld = [&](const double2* p) -> C { return reinterpret_cast<const C*>(p)[o]; };
    auto ldcf = [&](const double2* p) -> C {
      const CR z = reinterpret_cast<const CR*>(p)[o];
      return C{(double)z.x, (double)z.y};
    };
    auto stcf = [&](double2* p, C v) { reinterpret_cast<CR*>(p)[o] = CR{(R)v.x, (R)v.y}; };
    // the state of a field at plane j + d before any LES wall values (the
    // plane is uniform across the block)
    auto st = [&](const double2* full, const double2* wall, int d) -> C {
      const int pl = j + d;
      if (wall && (pl <= 1 || pl >= g.ny)) {
        const int slot = pl <= 1 ? pl : pl - g.ny + 2;
        return reinterpret_cast<const C*>(wall)[(long long)slot * g.sp + (o - (long long)j * g.sp)];
      }
      return reinterpret_cast<const C*>(full)[o + (long long)d * g.sp];
    };
    C cu1 = {0, 0}, cu3 = {0, 0};
    if (in1 || a.th1 > a.th0) {
      cu1 = st(a.cu1, a.wcu1, 0);
      cu3 = st(a.cu3, a.wcu3, 0);
    }
    if (in1) {
      C cp = reinterpret_cast<const C*>(a.cp)[o];
      C f1 = (-a.nu * visc) * cu1 + a.coriolis * cu3;
      C f3 = (-a.nu * visc) * cu3 - a.coriolis * cu1;
      for (int n = 0; n < a.nth; ++n) {
        C th = st(a.cth[n], a.wcth[n], 0);
        f1 = f1 + a.grav_x * th;
        f3 = f3 + a.grav_z * th;
      }
      if (a.fr_yf) {  // forcing.f90:55-72
        const double cy = a.fr_yf[j];
        const C u2m = 0.5 * (st(a.cu2, a.wcu2, 0) + st(a.cu2, a.wcu2, 1));
        f1 = f1 - (a.fr_sx * cy) * ik(kzk, cu1) + (a.fr_sz * cy) * ik(kxq, cu1) + (a.fr_sz * a.fr_dr) * u2m;
        f3 = f3 - (a.fr_sx * cy) * ik(kzk, cu3) + (a.fr_sz * cy) * ik(kxq, cu3) - (a.fr_sx * a.fr_dr) * u2m;
      }
      if (a.lesf1) {
        f1 = f1 + reinterpret_cast<const C*>(a.lesf1)[o];
        f3 = f3 + reinterpret_cast<const C*>(a.lesf3)[o];
      }
      f1 = f1 - (Z(0) + ik(kzk, Z(1)));
      f3 = f3 - (Z(2) + ik(kzk, Z(3)));
      C r1 = cu1, r3 = cu3;
      if (a.incr == 1) {
        r1 = r1 - ld(a.cub1);
        r3 = r3 - ld(a.cub3);
      } else if (a.incr == 2) {
        r1 = C{0.0, 0.0};
        r3 = C{0.0, 0.0};
      }
      if (a.rk_step > 1) {
        r1 = r1 + a.temp3 * ldcf(a.cf1);
        r3 = r3 + a.temp3 * ldcf(a.cf3);
      }
      r1 = r1 - a.temp4 * ik(kxq, cp);
      r3 = r3 - a.temp4 * ik(kzk, cp);
      if (mean) r1.x -= a.temp4 * a.force;
      r1 = r1 + a.temp5 * f1;
      r3 = r3 + a.temp5 * f3;
      stcf(a.cf1, f1);
      stcf(a.cf3, f3);
      put(0, r1);
      put(1, r3);
    }
    if (in2) {
      C cu2 = st(a.cu2, a.wcu2, 0);
      C cpj = reinterpret_cast<const C*>(a.cp)[o];
      C cpm = reinterpret_cast<const C*>(a.cp)[o - g.sp];
      C f2 = (-a.nu * visc) * cu2;
      if (!mean) {
        for (int n = 0; n < a.nth; ++n) {
          C th = st(a.cth[n], a.wcth[n], 0);
          C thm = st(a.cth[n], a.wcth[n], -1);
          f2 = f2 + (a.grav_y / (2.0 * cdy)) * (cdyfm * th + cdyf * thm);
        }
      }
      if (a.fr_y) {  // forcing.f90:74-84
        const double cy = a.fr_y[j];
        f2 = f2 - (a.fr_sx * cy) * ik(kzk, cu2) + (a.fr_sz * cy) * ik(kxq, cu2);
      }
      if (a.lesf2) f2 = f2 + reinterpret_cast<const C*>(a.lesf2)[o];
      f2 = f2 - (Z(4) + ik(kzk, Z(5)));
      C r2 = cu2;
      if (a.incr == 1) r2 = r2 - ld(a.cub2);
      if (a.incr == 2) r2 = C{0.0, 0.0};
      if (a.rk_step > 1) r2 = r2 + a.temp3 * ldcf(a.cf2);
      r2 = r2 - (a.temp4 / cdy) * (cpj - cpm);
      r2 = r2 + a.temp5 * f2;
      stcf(a.cf2, f2);
      put(2, r2);
    }
    for (int n = a.th0; n < a.th1; ++n) {
      if (!((inT >> n) & 1u)) continue;
      C th = st(a.cth[n], a.wcth[n], 0);
      C ft = (-(a.nu / a.pr[n]) * visc) * th;
      ft = ft - a.dthdx[n] * cu1 - a.dthdz[n] * cu3;
      if (a.fr_yf) {  // forcing.f90:87-96
        const double cy = a.fr_yf[j];
        ft = ft - (a.dthdx[n] * cy) * ik(kzk, th) + (a.dthdz[n] * cy) * ik(kxq, th);
      }
      if (a.lesfth[n]) ft = ft + reinterpret_cast<const C*>(a.lesfth[n])[o];
      ft = ft - (Z(6 + 2 * n) + ik(kzk, Z(7 + 2 * n)));
      C rt = th;
      if (a.incr == 1) rt = rt - ld(a.cthb[n]);
      if (a.incr == 2) rt = C{0.0, 0.0};
      if (a.rk_step > 1) rt = rt + a.temp3 * ldcf(a.cfth[n]);
      rt = rt + a.temp5 * ft;
      stcf(a.cfth[n], ft);
      put(3 + n, rt);
    }
  }
  if (PRE || spec) return;  // cr stays spectral
  // zero the dealiased gap of the cr slots (fft.f90:336-340)
  for (int lc = 0; lc < ncr; ++lc)
    for (int k = g.nkz + 1 + threadIdx.x; k < g.nz - g.nkz; k += blockDim.x)
      smem[lc * L + pidx(k)] = CR{(R)0, (R)0};
  // --- backward z transform of cr
  auto cr_used = [&](int c) -> bool {
    if (c < 2) return in1;
    if (c == 2) return in2;
    return (inT >> (c - 3)) & 1u;
  };
  for (int c0 = 0; c0 < ncr; c0 += ngroups) {
    const int lc = c0 + grp;
    const bool act = lc < ncr && cr_used(gcr(lc));
    CR* buf = smem + (long long)lc * L;
    __syncthreads();  // the cr slots were written by every thread of the block
    CR v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = buf[pidx(gt + G * r)];
    fft_any<R, 1, WAVE, LOGN>(v, buf, g.nz, g.lognz, twl, gt);
    if (act) {
      CR* out = reinterpret_cast<CR*>(a.Tcr[gcr(lc)]) + to;
      for (int k = gt; k < g.nz; k += G) out[k] = buf[pidx(k)];
    }
  }
}

// ------------------------------------------------------ tridiagonal solves --
// Real Crank-Nicolson solve of channel.f90:656-707 / :716-877 for one field,
// one (x,z) column per lane, rows j = 0..Ny+1.  Without LES the matrix rows
// depend only on j (dy, dyf, temp1, nu_v_scale, Pr), so the LU factors of the
// reference's Thomas sweep (thomas_forward_real_mpi / thomas_backward_real_mpi,
// :1335-1405, :1460-1496) are computed once on the host.  Each column's rows
// are split over S segments of L rows (one wavefront per segment); a segment
// runs its part of both sweeps in registers and the segments are joined by
// the lane-independent carry products (linear first-order recurrences).
// The workgroup index along x remapped so that consecutive indices run on one
// XCD.  The dispatcher deals workgroups to the 8 XCDs in turn (workgroup p on
// XCD p % 8); XCD x gets the contiguous range of logical indices from
// x q + min(x, r) on (n = 8 q + r), so tiles that straddle the same 128-byte
// lines -- the tridiagonal kernels' plane-strided row runs start at any
// 16-byte offset, a plane being sp modes -- read them through one L2.  At
// 512x257x512 the projection went 1.43 -> 1.32 ms per step with it and the
// spectral CN solves 2.17 -> 2.11 (mixed 2.05 -> 1.97).
__device__ __forceinline__ unsigned xcd_block(unsigned p, unsigned n) {
  const unsigned q = n / 8, r = n % 8, x = p % 8, i = p / 8;
  return x * q + (x < r ? x : r) + i;
}

struct SegRow {
  double m;    // forward multiplier  -a(j)/b'(j-1)            (0 at j = 0)
  double pm;   // prod of m from the segment's first row to j
  double e;    // 1/b'(j)
  double f;    // -c(j)/b'(j)
  double qf;   // prod of f from j to the segment's last row
  double wp, wm;  // explicit stencil weights (assembled rows)
  int kind;    // 1 = assembled row (vec = r + explicit), 0 = fixed row (vec = vb),
               // 2 = fixed row with vec = vb * (current time-dependent BC value),
               // 3 = plain row (vec = r; the transposed solves of the adjoint)
  double vb;   // fixed-row right-hand side
  double ra, rb, rc;  // the matrix row itself (used for fixed rows by the increment form)
};

// C adjacent columns per lane, loaded and stored as one vector of C values:
// a wavefront touches 64*C*sizeof(T) contiguous bytes of each row (the rows
// of one column are a plane apart, so short row accesses leave HBM idle).
// T is the storage type of u and TR that of r (the arithmetic is FP64).  INCR is the
// increment form: r holds the right-hand side less the state (RHS pass with
// incr), and the solve is for d = x - u with the same factors, A d = vec - A u:
// assembled rows (A = I - E, vec = r + u + E u) give r + 2 E u, fixed rows
// vec - (a u(j-1) + b u(j) + c u(j+1)).  u is overwritten with d.
//
// A column is any sequence of values a plane (pp) apart.  The physical solve
// has one per (x, z) point; the spectral solve (ncol = 2 sp, nbc = 1) one per
// real or imaginary part of a dealiased (kx, kz) mode, where only the mean
// mode's real part carries the wall values.  ncol = 0 means every launched
// column exists (the launchers always pass the column count), nbc = 0 that
// every column takes the wall values.
//
// ADJ (the adjoint's transposed solve vbar = A^-T xbar, rows of kind 3): the
// solution is not stored; instead rbar = vbar on the forward matrix's
// assembled rows (0 elsewhere) and ubar = E^T vbar, with E u = wp (u(j+1) -
// u(j)) - wm (u(j) - u(j-1)) from the forward row table prow, go to rbar and
// ubar for every row.
//
// The spectral transposed solve runs in place (r == ubar, the spectral state
// adjoint): input rows outside [ca, cb] (the planes the forward did not
// solve) are zero for the solve and pass through to ubar, which therefore
// holds the adjoint state's old value there plus E^T vbar.  Each row of a
// column is read, then written, by one lane, so r and ubar carry no
// __restrict__.
//
// CnFields batches the solves of several fields into one launch (blockIdx.y
// = field; nf = 0: the single field of the plain arguments): one launch
// keeps the CUs busy through the last blocks of each field.
constexpr int CN_MAXF = 3 + MAXTH;
template <typename T>
struct CnFields {
  const void* r[CN_MAXF];  // of the kernel's TR
  T* u[CN_MAXF];
  const SegRow* rows[CN_MAXF];
  int ca[CN_MAXF], cb[CN_MAXF];
  int nf;
};

template <int L, int C, typename T, bool INCR, bool ADJ = false, typename TR = T, bool RPLUS = false>
__global__ void __launch_bounds__(1024) k_thomas_seg(const TR* r,
                                                     T* __restrict__ u,
                                                     const SegRow* __restrict__ rows, int n, int S,
                                                     long long pp, int ca, int cb, double kscale,
                                                     const SegRow* __restrict__ prow = nullptr,
                                                     double* __restrict__ rbar = nullptr,
                                                     double* ubar = nullptr,
                                                     long long ncol = 0, long long nbc = 0,
                                                     CnFields<T> mf = CnFields<T>{}) {
  typedef T VN __attribute__((ext_vector_type(C)));  // C columns, one streaming access
  typedef TR VR __attribute__((ext_vector_type(C)));
  if (mf.nf > 0) {
    const int f = blockIdx.y;
    r = static_cast<const TR*>(mf.r[f]);
    u = mf.u[f];
    rows = mf.rows[f];
    ca = mf.ca[f];
    cb = mf.cb[f];
  }
  __shared__ double carry_f[16][64 * C];
  __shared__ double carry_b[16][64 * C];
  // The block's rows of the factor table (and, for the adjoint's epilogue,
  // of the forward table) staged in LDS once.  Read from the table itself,
  // each row's fields were scalar loads that the sweeps waited on one by one.
  constexpr bool SPROW = ADJ && L <= 32;  // LDS for both tables
  __shared__ SegRow srow[16 * L];
  __shared__ SegRow sprow[SPROW ? 16 * L : 1];
  {
    static_assert(sizeof(SegRow) % sizeof(double) == 0, "SegRow is copied as doubles");
    constexpr int W = sizeof(SegRow) / sizeof(double);
    const int nt = blockDim.x * blockDim.y, tid = threadIdx.y * blockDim.x + threadIdx.x;
    const double* src = reinterpret_cast<const double*>(rows);
    double* dst = reinterpret_cast<double*>(srow);
    for (int i = tid; i < n * W; i += nt) dst[i] = src[i];
    if constexpr (SPROW) {
      const double* ps = reinterpret_cast<const double*>(prow);
      double* pd = reinterpret_cast<double*>(sprow);
      for (int i = tid; i < n * W; i += nt) pd[i] = ps[i];
    }
  }
  __syncthreads();
  auto prow_at = [&](int jj) -> const SegRow& {
    if constexpr (SPROW)
      return sprow[jj];
    else
      return prow[jj];
  };
  // blockDim.x == 64: a wavefront is one segment, so seg is wave-uniform and
  // the row-table reads below become scalar loads
  const int lane = threadIdx.x, seg = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const long long col = ((long long)xcd_block(blockIdx.x, gridDim.x) * 64 + lane) * C;
  // lanes past the last column (the spectral solve's column count is not a
  // multiple of 64 C) take part in the carries but touch no memory; C divides
  // ncol, so a lane's columns are all in range or all out
  const bool inb = ncol == 0 || col < ncol;
  const int s0 = seg * L;
  double v[L][C];
  double w[L + 2][C];
  // u (the state, for the explicit term and the increment form) is not read
  // by the adjoint's transposed solve, whose rows are all plain (kind 3) and
  // which passes no u
#pragma unroll
  for (int t = 0; t < L + 2; ++t) {
    const int jj = s0 - 1 + t;
    if (!ADJ && inb && jj >= 0 && jj < n) {
      const VN x = __builtin_nontemporal_load(reinterpret_cast<const VN*>(u + (long long)jj * pp + col));
#pragma unroll
      for (int c = 0; c < C; ++c) w[t][c] = (double)x[c];
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) w[t][c] = 0.0;
    }
  }
  // The r rows go through a ring of RD loads in flight: the load of row t
  // + RD is issued before row t is assembled.  Each row's load used to sit
  // behind that row's kind test (a scalar load of the factor table), which
  // left one round trip to memory per row; loading every row up front
  // instead holds them all beside the u rows and spills.
  constexpr int RD = 4;
  auto ld_r = [&](int t) -> VR {
    const int jj = s0 + t;
    if (t < L && inb && jj < n)
      return __builtin_nontemporal_load(reinterpret_cast<const VR*>(r + (long long)jj * pp + col));
    return (VR)(TR)0;
  };
  VR rq[RD];
#pragma unroll
  for (int t = 0; t < RD; ++t) rq[t] = ld_r(t);
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int jj = s0 + t;
    const VR x = rq[t % RD];
    rq[t % RD] = ld_r(t + RD);
#pragma unroll
    for (int c = 0; c < C; ++c) v[t][c] = 0.0;
    if (jj < n) {
      const SegRow rw = srow[jj];
      if ((rw.kind == 1 || rw.kind == 3) && inb) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
          double val = (double)x[c];
          if (rw.kind == 1) {
            const double ex = rw.wp * (w[t + 2][c] - w[t + 1][c]) - rw.wm * (w[t + 1][c] - w[t][c]);
            val += INCR ? 2.0 * ex : ex;
            // RPLUS: r holds the right-hand side less the state (an FP32
            // increment); the state is added back here, in FP64
            if (RPLUS) val += w[t + 1][c];
          }
          if (ADJ && (jj < ca || jj > cb)) {  // passes through (w is free in ADJ)
            w[t + 1][c] = val;
            val = 0.0;
          }
          v[t][c] = val;
        }
      } else if (rw.kind == 0 || rw.kind == 2) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
          // the wall values are uniform in x and z: every column of the
          // physical solve, only the mean mode's real part (column 0) of the
          // spectral one (nbc = 1)
          double val = (nbc == 0 || col + c < nbc) ? ((rw.kind == 2) ? rw.vb * kscale : rw.vb) : 0.0;
          if (INCR) val -= rw.ra * w[t][c] + rw.rb * w[t + 1][c] + rw.rc * w[t + 2][c];
          v[t][c] = val;
        }
      }
    }
  }
  // forward sweep, local part
  double h[C];
#pragma unroll
  for (int c = 0; c < C; ++c) h[c] = 0.0;
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int jj = s0 + t;
    if (jj < n) {
      const double m = srow[jj].m;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        h[c] = v[t][c] + m * h[c];
        v[t][c] = h[c];
      }
    }
  }
  const int s1 = min(s0 + L, n) - 1;  // last row of this segment
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double last = 0.0;
#pragma unroll
    for (int t = 0; t < L; ++t)
      if (s0 + t == s1) last = v[t][c];
    carry_f[seg][lane * C + c] = last;
  }
  __syncthreads();
  double gin[C];
#pragma unroll
  for (int c = 0; c < C; ++c) gin[c] = 0.0;
  for (int sp = 0; sp < seg; ++sp) {
    const int l1 = min(sp * L + L, n) - 1;
    const double pm = srow[l1].pm;
#pragma unroll
    for (int c = 0; c < C; ++c) gin[c] = carry_f[sp][lane * C + c] + pm * gin[c];
  }
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int jj = s0 + t;
    if (jj < n) {
      const double pm = srow[jj].pm;
#pragma unroll
      for (int c = 0; c < C; ++c) v[t][c] += pm * gin[c];
    }
  }
  // backward sweep, local part
  double x[C];
#pragma unroll
  for (int c = 0; c < C; ++c) x[c] = 0.0;
#pragma unroll
  for (int t = L - 1; t >= 0; --t) {
    const int jj = s0 + t;
    if (jj < n) {
      const double e = srow[jj].e, f = srow[jj].f;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        x[c] = v[t][c] * e + f * x[c];
        v[t][c] = x[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) carry_b[seg][lane * C + c] = v[0][c];
  __syncthreads();
  double xin[C];
#pragma unroll
  for (int c = 0; c < C; ++c) xin[c] = 0.0;
  for (int sp = S - 1; sp > seg; --sp) {
    const double qf = srow[sp * L].qf;
#pragma unroll
    for (int c = 0; c < C; ++c) xin[c] = carry_b[sp][lane * C + c] + qf * xin[c];
  }
  if constexpr (ADJ) {
    // the solution of every row of the segment, then its neighbours' edge rows
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int jj = s0 + t;
      const double qf = srow[jj < n ? jj : n - 1].qf;
#pragma unroll
      for (int c = 0; c < C; ++c) v[t][c] += qf * xin[c];
    }
    __syncthreads();  // every lane is done with the carries
#pragma unroll
    for (int c = 0; c < C; ++c) {
      double last = 0.0;
#pragma unroll
      for (int t = 0; t < L; ++t)
        if (s0 + t == s1) last = v[t][c];
      carry_f[seg][lane * C + c] = last;     // row s1 of this segment
      carry_b[seg][lane * C + c] = v[0][c];  // row s0
    }
    __syncthreads();
    double vm[C], vp[C];  // rows s0 - 1 and s1 + 1
#pragma unroll
    for (int c = 0; c < C; ++c) {
      vm[c] = seg > 0 ? carry_f[seg - 1][lane * C + c] : 0.0;
      vp[c] = seg + 1 < S ? carry_b[seg + 1][lane * C + c] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int jj = s0 + t;
      if (jj < n) {
        const SegRow rj = prow_at(jj);
        const bool km = jj >= 1 && prow_at(jj - 1).kind == 1, kp = jj + 1 < n && prow_at(jj + 1).kind == 1;
        const double wpm = km ? prow_at(jj - 1).wp : 0.0, wmp = kp ? prow_at(jj + 1).wm : 0.0;
        typedef double VD __attribute__((ext_vector_type(C)));
        VD rb, ub;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double x = v[t][c];
          const double xm = (t == 0) ? vm[c] : v[t > 0 ? t - 1 : 0][c];
          const double xp = (jj == s1) ? vp[c] : v[t + 1 < L ? t + 1 : t][c];
          double b = wpm * xm + wmp * xp;
          if (rj.kind == 1) b -= (rj.wp + rj.wm) * x;
          ub[c] = b + w[t + 1][c];
          rb[c] = (rj.kind == 1) ? x : 0.0;
        }
        const long long o = (long long)jj * pp + col;
        if (inb) {
          __builtin_nontemporal_store(rb, reinterpret_cast<VD*>(rbar + o));
          __builtin_nontemporal_store(ub, reinterpret_cast<VD*>(ubar + o));
        }
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int jj = s0 + t;
      if (inb && jj < n && jj >= ca && jj <= cb) {
        const double qf = srow[jj].qf;
        VN o;
#pragma unroll
        for (int c = 0; c < C; ++c) o[c] = (T)(v[t][c] + qf * xin[c]);
        __builtin_nontemporal_store(o, reinterpret_cast<VN*>(u + (long long)jj * pp + col));
      }
    }
  }
}

// The kx = 0 column of a spectral field on planes [ca, cb], projected as the
// physical round trip F P projects it (FFTW's c2r reads the real part of the
// kx = 0 line): c(kz) -> (c(kz) + conj c(-kz)) / 2, Im c(0) -> 0.  The spectral
// CN solve applies it to its solution (the reference's solution is the
// transform of a real field), the spectral transposed solve to its input; for
// a state that already has the symmetry it changes nothing.
__global__ void k_herm0(double2* __restrict__ c, Geo g, int ca, int cb) {
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nkz
  const int j = ca + blockIdx.y;
  if (kk > g.nkz || j > cb) return;
  double2* row = c + (long long)j * g.sp;  // q = 0
  if (kk == 0) {
    row[0].y = 0.0;
  } else {
    const double2 a = row[kk], b = row[g.nkk - kk];
    const double2 s = {0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
    row[kk] = s;
    row[g.nkk - kk] = double2{s.x, -s.y};
  }
}

// ----------------------------------------------------------- projection --
// rem_div_chan (channel.f90:946-1023) with apply_BC_rem_div_mpi (:1974-2015)
// and the complex Thomas sweeps (:1408-1457, :1499-1540); optionally the
// pressure update cp += phi/h (:905-911).
//
// A mode's matrix depends only on lam = kx^2 + kz^2 and the grid, so its
// pivots b_j (the forward elimination of the Thomas sweep) are computed once
// per handle by k_remdiv_pivots and stored as 1/b_j.  The solve is then two
// first-order linear recurrences per mode,
//   g_j   = rhs_j + m_j g_{j-1},           m_j = -ml_j / b_{j-1}    (forward)
//   phi_j = g_j / b_j - (mu_j / b_j) phi_{j+1}                      (backward)
// which k_remdiv_seg splits into segments of L rows held in registers and
// joins with a scan over the segments.
struct RemDivArgs {
  double2* cu1;
  double2* cu2;
  double2* cu3;
  double2* cp;
  const double* ib;     // 1/b_j per (row, mode), rows 1..Ny (ny2 x sp)
  const double* kx;
  const double* kz;
  double inv_h;  // 1/temp4
  int update_cp;
  int js, je;
};

// Row coefficients of the projection's tridiagonal (channel.f90:970-976,
// :1988-2011): rows 2..Ny-1 ml = 1/(dy(j) dyf(j)), mu = 1/(dy(j+1) dyf(j)),
// md = -lam - mu - ml; row Ny is phi(Ny) = phi(Ny-1) (ml = 1, md = -1,
// mu = 0); row 1 has b = 1 and mu = -1 (Neumann) or 0 (Dirichlet, the (0,0)
// mode), set per mode in the kernels.  idyf, idy: 1/dyf(j), 1/dy(j).
struct PoisRow {
  double ml, mu, idyf, idy;
};

// The pivots depend on kx^2 + kz^2, so a mode and its -kz partner share them:
// the table holds kz >= 0 only, nq (Nkz+1) entries a row (piv_plane), and a
// mode (q, kk) reads entry piv_index (kk > Nkz is -kz of 2Nkz+1-kk).
__host__ __device__ inline long long piv_plane(const Geo& g) { return (long long)g.nq * (g.nkz + 1); }
__device__ inline long long piv_index(const Geo& g, int q, int kk) {
  return (long long)q * (g.nkz + 1) + (kk <= g.nkz ? kk : g.nkk - kk);
}

__global__ void __launch_bounds__(256) k_remdiv_pivots(double* __restrict__ ib, const PoisRow* __restrict__ rows,
                                                       const double* __restrict__ kx,
                                                       const double* __restrict__ kz, Geo g) {
  const long long mh = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long hp = piv_plane(g);
  if (mh >= hp) return;
  const int q = (int)((unsigned)mh / (unsigned)(g.nkz + 1)), kk = (int)((unsigned)mh % (unsigned)(g.nkz + 1));
  const double lam = kx[q] * kx[q] + kz[kk] * kz[kk];
  double bprev = 1.0;
  double muprev = (mh == 0) ? 0.0 : -1.0;  // the (0,0) mode's Dirichlet row 1
  ib[hp + mh] = 1.0;
  for (int j = 2; j <= g.ny; ++j) {
    const PoisRow rw = rows[j];
    const double md = (j < g.ny) ? -lam - rw.mu - rw.ml : -1.0;
    const double mlt = -rw.ml / bprev;
    const double b = md + mlt * muprev;
    ib[(long long)j * hp + mh] = 1.0 / b;
    bprev = b;
    muprev = rw.mu;
  }
}

// One workgroup: MW consecutive modes x S segments of L rows.  Row 1 needs
// no segment (g_1 = 0, and phi_1 = phi_2, or 0 for the (0,0) mode), so the
// segments cover rows 2..Ny: segment s holds rows 2 + s L ... .  Lane
// mm + MW sl of wavefront w is mode mm of segment s = SPW w + sl, SPW = 64/MW
// segments per wavefront, so a wavefront reads MW-mode runs of SPW L rows
// (the wider the runs, the better HBM serves these plane-strided reads:
// 8-mode runs measured 1.94 ms per step at 512x257x512, 16-mode runs 1.66).
// Every value of a segment stays in registers between the load and the
// store.  The segments are joined by an inclusive scan over s (forward: g at
// each segment's last row) and a reverse one (backward: phi at each
// segment's first row): within a wavefront by shuffles, across wavefronts by
// one exchange of the wavefront totals in LDS.
// MAXT bounds the workgroup (and so the registers per lane): MW * S <= MAXT.
template <int L, int MW, int MAXT>
__global__ void __launch_bounds__(MAXT) k_remdiv_seg(RemDivArgs a, Geo g, const PoisRow* __restrict__ rows,
                                                     int S) {
  typedef cpx<double> C;
  constexpr int SPW = 64 / MW;
  constexpr int SMAX = MAXT / MW, WMAX = MAXT / 64;
  __shared__ double tP[WMAX][MW];
  __shared__ C tA[WMAX][MW];
  __shared__ double tQ[WMAX][MW];
  __shared__ C tB[WMAX][MW];
  __shared__ C tL[SMAX][MW];
  // the row coefficients of the planes the block's segments cover, staged in
  // LDS once (read from the table, the lanes of a wavefront's segments read
  // different rows in every sweep step)
  __shared__ PoisRow srw[2 + L * SMAX];
  for (int i = threadIdx.x; i <= g.ny && i < 2 + L * SMAX; i += blockDim.x) srw[i] = rows[i];
  __syncthreads();
  const int mm = threadIdx.x % MW, s = threadIdx.x / MW;
  const int w = threadIdx.x / 64, sl = (threadIdx.x % 64) / MW;
  const int nw = blockDim.x / 64;
  const long long sp = g.sp;
  const long long m = (long long)xcd_block(blockIdx.x, gridDim.x) * MW + mm;
  const bool live = m < sp;
  const long long mc = live ? m : 0;
  const int q = (int)((unsigned)mc / (unsigned)g.nkk), kk = (int)((unsigned)mc % (unsigned)g.nkk);
  const double kxq = a.kx[q], kzk = a.kz[kk];
  const int ny = g.ny;
  const int j0 = 2 + s * L;
  const C Z = {0.0, 0.0};
  // a field is below 4 GB (checked by the launcher): one 32-bit byte offset
  // per row serves every field, from the field's base in scalar registers
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
  // the pivot of this mode at row j (shared with its -kz partner)
  const long long hp = piv_plane(g), mh = piv_index(g, q, kk);
  auto offp = [&](int j) { return (unsigned)(((long long)j * hp + mh) * 8); };
  auto shup = [](C v, int d) { return C{__shfl_up(v.x, d), __shfl_up(v.y, d)}; };
  auto shdn = [](C v, int d) { return C{__shfl_down(v.x, d), __shfl_down(v.y, d)}; };
  // rows j0 .. j0+L-1 of cu1, cu3, cu2 (plus cu2 at j0+L) and 1/b (plus j0-1)
  C d1[L], d3[L], d2[L + 1];
  double ibv[L + 1];
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    const unsigned o = off(j);
    const bool ok = live && j <= ny;
    d1[t] = ok ? ldc(a.cu1, o) : Z;
    d3[t] = ok ? ldc(a.cu3, o) : Z;
    d2[t] = ok ? ldc(a.cu2, o) : Z;
    ibv[t + 1] = ok ? ldd(a.ib, offp(j)) : 0.0;
  }
  d2[L] = (live && j0 + L <= ny) ? ldc(a.cu2, off(j0 + L)) : Z;
  ibv[0] = (live && j0 - 1 <= ny) ? ldd(a.ib, offp(j0 - 1)) : 0.0;

  // forward sweep of the segment from a zero carry: g_t = gl_t + Pl_t * g(j0-1)
  C gl[L];
  double Pl[L];
  {
    C h = Z;
    double P = 1.0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
      const int j = j0 + t;
      if (j <= ny) {
        const PoisRow rw = srw[j];
        const double mj = -rw.ml * ibv[t];
        const C rhs = (j < ny) ? ik(kxq, d1[t]) + rw.idyf * (d2[t + 1] - d2[t]) + ik(kzk, d3[t]) : Z;
        h = rhs + mj * h;
        P = mj * P;
      }
      gl[t] = h;
      Pl[t] = P;
    }
  }
  // inclusive scan of (P, A) over the segments: in the wavefront, then across
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
    // lane sl-1's inclusive value: the composition of the segments before
    // this one in the wavefront
    double Pe = __shfl_up(P, MW);
    C Ae = shup(A, MW);
    if (sl == 0) {
      Pe = 1.0;
      Ae = Z;
    }
    __syncthreads();
    C cin = Z;  // g at the last row of wavefront w-1 (g_1 = 0 before wavefront 0)
    for (int w2 = 0; w2 < w; ++w2) cin = tA[w2][mm] + tP[w2][mm] * cin;
    const C gin = Ae + Pe * cin;
#pragma unroll
    for (int t = 0; t < L; ++t) gl[t] = gl[t] + Pl[t] * gin;
  }
  // backward sweep of the segment from a zero carry: phi_t = gl_t + Pl_t * phi(j0+L)
  {
    C x = Z;
    double Q = 1.0;
#pragma unroll
    for (int t = L - 1; t >= 0; --t) {
      const int j = j0 + t;
      if (j <= ny) {
        const double e = ibv[t + 1];
        const double f = -srw[j].mu * e;  // mu(Ny) = 0
        x = e * gl[t] + f * x;
        Q = f * Q;
      }
      gl[t] = x;
      Pl[t] = Q;
    }
  }
  // cp is read here, so that its latency overlaps the reverse scan
  const bool cpu = a.update_cp != 0;
  C pv[L];
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    pv[t] = (live && cpu && j <= ny && j >= a.js && j <= a.je) ? ldc(a.cp, off(j)) : Z;
  }
  // reverse inclusive scan of (Q, B): phi at each segment's first row
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
    C ctop = Z;  // phi at the first row of wavefront w+1 (0 above the last)
    for (int w2 = nw - 1; w2 > w; --w2) ctop = tB[w2][mm] + tQ[w2][mm] * ctop;
    const C ptop = Be + Qe * ctop;
#pragma unroll
    for (int t = 0; t < L; ++t) gl[t] = gl[t] + Pl[t] * ptop;  // gl now holds phi
  }
  // row 1: phi_1 = (g_1 - mu_1 phi_2) / b_1 = phi_2, or 0 for the (0,0) mode
  const C phi1 = (mc == 0) ? Z : gl[0];
  // phi at the row below the segment, for the cu2 update of its first row
  tL[s][mm] = gl[L - 1];
  __syncthreads();
  const C phim = (s > 0) ? tL[s - 1][mm] : phi1;
  if (!live) return;
#pragma unroll
  for (int t = 0; t < L; ++t) {
    const int j = j0 + t;
    if (j > ny) break;
    const unsigned o = off(j);
    const C phi = gl[t];
    stc(a.cu1, o, d1[t] - ik(kxq, phi));
    stc(a.cu3, o, d3[t] - ik(kzk, phi));
    stc(a.cu2, o, d2[t] - srw[j].idy * (phi - (t == 0 ? phim : gl[t - 1])));
    if (cpu && j >= a.js && j <= a.je) stc(a.cp, o, pv[t] + a.inv_h * phi);
  }
  if (s == 0) {  // row 1: cu1, cu3 and cp (cu2 at row 1 is not part of the projection)
    const unsigned o = off(1);
    const C v1 = ldc(a.cu1, o), v3 = ldc(a.cu3, o);
    const bool cp1 = cpu && 1 >= a.js && 1 <= a.je;
    const C vp = cp1 ? ldc(a.cp, o) : Z;
    stc(a.cu1, o, v1 - ik(kxq, phi1));
    stc(a.cu3, o, v3 - ik(kzk, phi1));
    if (cp1) stc(a.cp, o, vp + a.inv_h * phi1);
  }
}

// Device-to-device copy of n complex values (checkpoints, the LES pre-BC
// state copy): a streaming kernel, four values per thread per iteration.
__global__ void __launch_bounds__(256) k_copy2(double2* __restrict__ dst, const double2* __restrict__ src,
                                               long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Planes 0, 1, Ny, Ny+1 of nf spectral fields into 4-plane copies (the
// state before the LES stage's wall values, read back by the RHS pass).
struct WallCopy {
  const double2* src[3 + MAXTH];
  double2* dst[3 + MAXTH];
  int nf;
};
__global__ void __launch_bounds__(256) k_copy_walls(WallCopy w, Geo g) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.sp) return;
  const int slot = blockIdx.y, f = blockIdx.z;
  const int pl = slot <= 1 ? slot : g.ny - 2 + slot;
  w.dst[f][(long long)slot * g.sp + m] = w.src[f][(long long)pl * g.sp + m];
}

// ---------------------------------------------------------- wall BCs ----
// apply_BC_vel_lower_post / apply_BC_vel_upper_post (boundary.f90:299-498).
struct BcPostArgs {
  double2* cu1;
  double2* cu2;
  double2* cu3;
  const double* dy;
  const double* dyf;
  int u_lo, v_lo, w_lo, u_hi, v_hi, w_hi;
  double u_lo_c1, v_lo_c1, w_lo_c1, u_hi_c1, v_hi_c1, w_hi_c1;
  double w_hi_c1_transient;
};

__global__ void k_bc_post(BcPostArgs a, Geo g) {
  typedef cpx<double> C;
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.sp) return;
  const bool mean = (m == 0);
  const long long P = g.sp;
  const int ny = g.ny;
  C* cu1 = reinterpret_cast<C*>(a.cu1);
  C* cu2 = reinterpret_cast<C*>(a.cu2);
  C* cu3 = reinterpret_cast<C*>(a.cu3);
  const C Z = {0.0, 0.0};
  // lower wall
  if (a.u_lo == 0) {
    cu1[m] = Z;
    cu1[P + m] = Z;
    if (mean) cu1[P + m] = cu1[m] = C{a.u_lo_c1, 0.0};
  } else {
    cu1[P + m] = cu1[2 * P + m];
    cu1[m] = cu1[P + m];
    if (mean) {
      cu1[P + m].x = cu1[2 * P + m].x - a.dy[2] * a.u_lo_c1;
      cu1[m] = C{cu1[P + m].x - a.dy[1] * a.u_lo_c1, cu1[P + m].y};
    }
  }
  if (a.w_lo == 0) {
    cu3[m] = Z;
    cu3[P + m] = Z;
    if (mean) cu3[P + m] = cu3[m] = C{a.w_lo_c1, 0.0};
  } else {
    cu3[P + m] = cu3[2 * P + m];
    cu3[m] = cu3[P + m];
    if (mean) {
      cu3[P + m].x = cu3[2 * P + m].x - a.dy[2] * a.w_lo_c1;
      cu3[m] = C{cu3[P + m].x - a.dy[1] * a.w_lo_c1, cu3[P + m].y};
    }
  }
  if (a.v_lo == 0) {
    C t = cu2[2 * P + m];
    cu2[P + m] = C{2.0 * a.v_lo_c1 - t.x, -t.y};
    cu2[m] = cu2[P + m];
  } else {
    cu2[P + m] = cu2[2 * P + m];
    cu2[m] = cu2[P + m];
    if (mean) {
      cu2[P + m].x = cu2[2 * P + m].x - a.dyf[1] * a.v_lo_c1;
      cu2[m] = C{cu2[P + m].x - a.dyf[1] * a.v_lo_c1, cu2[P + m].y};
    }
  }
  // upper wall
  const long long T0 = (long long)ny * P + m, T1 = T0 + P, Tm = T0 - P;
  if (a.u_hi == 0) {
    cu1[T0] = Z;
    cu1[T1] = Z;
    if (mean) cu1[T0] = cu1[T1] = C{a.u_hi_c1, 0.0};
  } else {
    cu1[T0] = cu1[Tm];
    cu1[T1] = cu1[T0];
    if (mean) {
      cu1[T0].x = cu1[Tm].x + a.dy[ny] * a.u_hi_c1;
      cu1[T1] = C{cu1[T0].x + a.dy[ny] * a.u_hi_c1, cu1[T0].y};
    }
  }
  if (a.w_hi == 0) {
    cu3[T0] = Z;
    cu3[T1] = Z;
    if (mean) {
      cu3[T0] = C{a.w_hi_c1, 0.0};
      cu3[T1] = Z;  // boundary.f90:454 overwrites the ghost
    }
  } else {
    cu3[T0] = cu3[Tm];
    cu3[T1] = cu3[T0];
    if (mean) {
      cu3[T0].x = cu3[Tm].x + a.dy[ny] * a.w_hi_c1_transient;
      cu3[T1] = C{cu3[T0].x + a.dy[ny] * a.w_hi_c1_transient, cu3[T0].y};
    }
  }
  if (a.v_hi == 0) {
    if (mean) {
      C t = cu2[T0];
      cu2[T1] = C{2.0 * a.v_hi_c1 - t.x, -t.y};
    }
  } else {
    cu2[T0] = cu2[Tm];
    cu2[T1] = cu2[T0];
    if (mean) cu2[T1].x = cu2[T0].x + a.dy[ny] * a.v_hi_c1;
  }
}

}  // namespace dgpu
