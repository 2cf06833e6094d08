// dgpu_fft.h -- device-side FFT building blocks for gfx950 (CDNA4).
//
// The reference transforms every y-plane with FFTW guru plans (fft.f90:185-209):
// a real-to-complex FFT of length Nx along x, an MPI_Alltoall, then a complex FFT
// of length Nz along z.  Here the whole field sits in HBM, so each direction is
// one kernel pass over the field and the transpose is the pass's own store
// pattern.  The transforms themselves run in LDS: a Stockham autosort FFT
// (radix 8, then one radix-4 or radix-2 pass when log2 N is not a multiple of
// 3) executed by N/8 threads per transform, 8 complex values per thread in
// registers between passes.
//
// Conventions match FFTW's: forward = exp(-2 pi i k n / N), unnormalised;
// backward = exp(+2 pi i k n / N), unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace dgpu {

template <typename R>
struct cpx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cpx<R> operator*(R s, cpx<R> a) {
  return {s * a.x, s * a.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> conjg(cpx<R> a) {
  return {a.x, -a.y};
}
// i*k*a : the reference's cikx(i)*a with cikx = (0, kx)  (fft.f90:217, 229)
template <typename R>
__device__ __forceinline__ cpx<R> ik(R k, cpx<R> a) {
  return {-k * a.y, k * a.x};
}

// Store of a result that the next kernel, not this one, reads back: a
// streaming (nontemporal) vector store, which keeps the pass's output from
// displacing its own inputs in the caches (a plain-store copy reaches 4.6-4.8
// TB/s on MI355X at full occupancy, a nontemporal one 5.2-6.0 TB/s).
template <typename R>
__device__ __forceinline__ void st_out(cpx<R>* p, cpx<R> v) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
}
template <typename R>
__device__ __forceinline__ void st_out(R* p, R v) {
  __builtin_nontemporal_store(v, p);
}
// Load of an input that this pass reads exactly once (streaming load).
template <typename R>
__device__ __forceinline__ R ld_in(const R* p) {
  return __builtin_nontemporal_load(p);
}
template <typename R>
__device__ __forceinline__ cpx<R> ld_in(const cpx<R>* p) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return {w.x, w.y};
}

// Small DFTs in registers.  S = -1 forward, +1 backward.
template <typename R, int S>
__device__ __forceinline__ void dft2(cpx<R>* v) {
  cpx<R> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <typename R, int S>
__device__ __forceinline__ void dft4(cpx<R>& v0, cpx<R>& v1, cpx<R>& v2, cpx<R>& v3) {
  cpx<R> a = v0 + v2, b = v0 - v2, c = v1 + v3, t = v1 - v3;
  cpx<R> d = {-(R)S * t.y, (R)S * t.x};  // (S i) * t
  v0 = a + c;
  v1 = b + d;
  v2 = a - c;
  v3 = b - d;
}

template <typename R, int S>
__device__ __forceinline__ void dft8(cpx<R>* v) {
  const R h = (R)0.70710678118654752440084436210485;
  cpx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cpx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<R, S>(e0, e1, e2, e3);
  dft4<R, S>(o0, o1, o2, o3);
  // twiddles w8^m, w8 = exp(S i pi/4)
  cpx<R> t1 = {h * (o1.x - (R)S * o1.y), h * (o1.y + (R)S * o1.x)};
  cpx<R> t2 = {-(R)S * o2.y, (R)S * o2.x};
  cpx<R> t3 = {h * (-o3.x - (R)S * o3.y), h * (-o3.y + (R)S * o3.x)};
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + t1;
  v[5] = e1 - t1;
  v[2] = e2 + t2;
  v[6] = e2 - t2;
  v[3] = e3 + t3;
  v[7] = e3 - t3;
}

// LDS index of element i of a transform buffer: one pad slot after every 8
// elements, so that the stride-8 stores of a radix-8 pass (and the strided
// loads that follow) fall in distinct LDS banks.  A buffer of N elements
// occupies fft_lds_len(N) slots.
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int fft_lds_len(int n) { return n + (n >> 3); }

// Distance between the transform buffers of the groups of an x pass.  The x
// passes also access the buffers across groups (consecutive lanes, the same
// element of consecutive groups' buffers); a stride of fft_lds_len(N), a
// multiple of the bank count, would put all those lanes in the same banks, so
// the buffers are spaced one element further apart.
__host__ __device__ constexpr int x_group_stride(int n) { return fft_lds_len(n) + 1; }

// Synchronisation between the passes of one transform.  When the G = N/8
// lanes of a transform lie inside one wavefront (G <= 64, groups aligned to
// G), LDS operations of the wavefront execute in program order, so a
// wavefront-scope fence plus a scheduling barrier is enough; larger
// transforms span wavefronts and need the workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// One Stockham pass of radix 2^LOGR over a length-N transform in LDS, after
// the radices already applied multiply to Ns = 2^logNs.  Each of the G = N/8
// threads owns 8 complex values (8 >> LOGR butterflies).
template <typename R, int S, int LOGR, bool WAVE>
__device__ __forceinline__ void stockham_pass(cpx<R>* buf, int N, int logNs,
                                              const cpx<R>* __restrict__ tw, int gt) {
  constexpr int RDX = 1 << LOGR, NB = 8 >> LOGR;
  const int G = N >> 3;
  const int stride = N >> LOGR;
  const int Ns = 1 << logNs;
  const int tstep = N >> (logNs + LOGR);
  cpx<R> v[8];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int r = 0; r < RDX; ++r) v[bb * RDX + r] = buf[pidx(gt + bb * G + r * stride)];
  fft_sync<WAVE>();
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) {
    const int j = gt + bb * G;
    const int k = j & (Ns - 1);
    if (logNs > 0) {
#pragma unroll
      for (int r = 1; r < RDX; ++r) {
        cpx<R> t = tw[k * r * tstep];
        if (S > 0) t.y = -t.y;
        v[bb * RDX + r] = v[bb * RDX + r] * t;
      }
    }
    if constexpr (LOGR == 3)
      dft8<R, S>(v);
    else if constexpr (LOGR == 2)
      dft4<R, S>(v[bb * 4], v[bb * 4 + 1], v[bb * 4 + 2], v[bb * 4 + 3]);
    else
      dft2<R, S>(v + bb * 2);
    const int idxD = ((j >> logNs) << (logNs + LOGR)) + k;
#pragma unroll
    for (int r = 0; r < RDX; ++r) buf[pidx(idxD + r * Ns)] = v[bb * RDX + r];
  }
  fft_sync<WAVE>();
}

// Complex FFT of length N (power of two, 8 <= N <= 4096) by the G = N/8
// threads of one group (gt = 0..G-1).  Input in registers: v[r] = x[gt + G r],
// the operands of the first radix-8 butterfly, which therefore runs without
// touching LDS.  Output in LDS at `buf` in natural order (padded, see pidx).
// tw[t] = exp(-2 pi i t / N).  The caller must not touch `buf` of this group
// between its last use and this call returning, except through the returned
// data.  Every thread of the workgroup calls this the same number of times
// with the same N when WAVE is false (workgroup barriers inside).
template <typename R, int S, bool WAVE>
__device__ __forceinline__ void fft_regs(cpx<R> (&v)[8], cpx<R>* buf, int N, int logN,
                                         const cpx<R>* __restrict__ tw, int gt) {
  dft8<R, S>(v);  // Ns = 1: no twiddles; outputs y[8 gt + r]
  fft_sync<WAVE>();  // the previous use of buf is finished
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[pidx(gt * 8 + r)] = v[r];
  fft_sync<WAVE>();
  int logNs = 3;
  while (logNs < logN) {
    const int rem = logN - logNs;
    if (rem >= 3) {
      stockham_pass<R, S, 3, WAVE>(buf, N, logNs, tw, gt);
      logNs += 3;
    } else if (rem == 2) {
      stockham_pass<R, S, 2, WAVE>(buf, N, logNs, tw, gt);
      logNs += 2;
    } else {
      stockham_pass<R, S, 1, WAVE>(buf, N, logNs, tw, gt);
      logNs += 1;
    }
  }
}

// The same transform with N = 2^LOGN a compile-time constant (N >= 64, so
// that the group size G = N/8 and every pass's stride are multiples of 8):
// each LDS index of a pass is then one per-thread base plus an immediate
// offset, pidx(gt + c) = pidx(gt) + c + c/8 for c a multiple of 8.
template <typename R, int S, int LOGR, bool WAVE, int LOGN, int LOGNS>
__device__ __forceinline__ void stockham_pass_c(cpx<R>* buf, const cpx<R>* __restrict__ tw, int gt) {
  constexpr int N = 1 << LOGN, RDX = 1 << LOGR, NB = 8 >> LOGR;
  constexpr int G = N >> 3, stride = N >> LOGR, Ns = 1 << LOGNS;
  constexpr int tstep = N >> (LOGNS + LOGR);
  static_assert(G % 8 == 0 && stride % 8 == 0, "compile-time passes need N >= 64");
  const int pg = pidx(gt);
  cpx<R> v[8];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int r = 0; r < RDX; ++r) {
      const int c = bb * G + r * stride;
      v[bb * RDX + r] = buf[pg + c + (c >> 3)];
    }
  fft_sync<WAVE>();
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) {
    const int j = gt + bb * G;
    const int k = j & (Ns - 1);
    if constexpr (LOGNS > 0) {
#pragma unroll
      for (int r = 1; r < RDX; ++r) {
        cpx<R> t = tw[k * r * tstep];
        if (S > 0) t.y = -t.y;
        v[bb * RDX + r] = v[bb * RDX + r] * t;
      }
    }
    if constexpr (LOGR == 3)
      dft8<R, S>(v);
    else if constexpr (LOGR == 2)
      dft4<R, S>(v[bb * 4], v[bb * 4 + 1], v[bb * 4 + 2], v[bb * 4 + 3]);
    else
      dft2<R, S>(v + bb * 2);
    const int idxD = ((j >> LOGNS) << (LOGNS + LOGR)) + k;
    const int pd = pidx(idxD);
#pragma unroll
    for (int r = 0; r < RDX; ++r) {
      // idxD + r Ns: Ns >= 8 keeps the offset a multiple of 8 (the first
      // pass, Ns = 1, writes consecutive slots instead)
      if constexpr (Ns >= 8)
        buf[pd + r * Ns + ((r * Ns) >> 3)] = v[bb * RDX + r];
      else
        buf[pidx(idxD + r * Ns)] = v[bb * RDX + r];
    }
  }
  fft_sync<WAVE>();
}

template <typename R, int S, bool WAVE, int LOGN, int LOGNS>
__device__ __forceinline__ void fft_passes_c(cpx<R>* buf, const cpx<R>* __restrict__ tw, int gt) {
  if constexpr (LOGNS < LOGN) {
    constexpr int REM = LOGN - LOGNS;
    constexpr int LR = REM >= 3 ? 3 : REM;
    stockham_pass_c<R, S, LR, WAVE, LOGN, LOGNS>(buf, tw, gt);
    fft_passes_c<R, S, WAVE, LOGN, LOGNS + LR>(buf, tw, gt);
  }
}

// fft_regs for N = 2^LOGN known at compile time.
template <typename R, int S, bool WAVE, int LOGN>
__device__ __forceinline__ void fft_regs_c(cpx<R> (&v)[8], cpx<R>* buf, const cpx<R>* __restrict__ tw, int gt) {
  dft8<R, S>(v);
  fft_sync<WAVE>();
  const int base = pidx(gt * 8);  // gt * 8 + r: 9 gt + r
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[base + r] = v[r];
  fft_sync<WAVE>();
  fft_passes_c<R, S, WAVE, LOGN, 3>(buf, tw, gt);
}

// ------------------------------------------ 512 points in one wavefront --
// A 512-point transform held by the 64 lanes of one wavefront, 8 values a
// lane, with the radix-8 steps on registers and the data exchanges between
// them done partly across lanes.  Index n = n0 + 8 n1 + 64 n2 with n0, n1 the
// low and high three bits of the lane and n2 the register; output
// k = k2 + 8 k1 + 64 k0:
//   X = sum_n0 w^(64 n0 k0) w^(8 n0 k1) w^(n0 k2)
//         sum_n1 w^(64 n1 k1) w^(8 n1 k2) sum_n2 w^(64 n2 k2) x
// (w = exp(S 2 pi i / 512)).  Step A is a radix-8 DFT over the registers,
// then the twiddle w^((n0 + 8 n1) k2) = w^(lane k2).  The high lane bits (n1)
// and the register index (k2) are then transposed without LDS: lane bit 5
// with v_permlane32_swap, bit 4 with v_permlane16_swap, bit 3 with two DPP
// moves (row_ror:8, each writing one half of every row through its bank
// mask).  Step B is a radix-8 DFT over n1, then the twiddle w^(8 n0 k1).  The
// low lane bits and the register index go through LDS (one write and one
// read a value), step C transforms over n0, and the result is stored in
// natural order, k = k2 + 8 k1 + 64 k0 from lane k1 + 8 k2, register k0.
// Against the Stockham form (fft_regs_c) this replaces one of the two LDS
// exchanges of the transform by 64 cross-lane moves per complex FP64 value
// set: 24 LDS accesses a value instead of 40.

// The 2x2 transposes of a lane bit and a register bit on 32-bit words: the
// values of (lane bit set, a) and (lane bit clear, b) trade places.
__device__ __forceinline__ void xl_bit5(unsigned& a, unsigned& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
__device__ __forceinline__ void xl_bit4(unsigned& a, unsigned& b) {
  const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
__device__ __forceinline__ void xl_bit3(unsigned& a, unsigned& b) {
  // row_ror:8 (DPP control 0x128) reads lane l ^ 8 within each 16-lane row;
  // bank mask 0xC writes lanes 8-15 of a row, 0x3 lanes 0-7
  const int na = __builtin_amdgcn_update_dpp((int)a, (int)b, 0x128, 0xF, 0xC, false);
  const int nb = __builtin_amdgcn_update_dpp((int)b, (int)a, 0x128, 0xF, 0x3, false);
  a = (unsigned)na;
  b = (unsigned)nb;
}

template <int BIT>
__device__ __forceinline__ void xl_word(unsigned& a, unsigned& b) {
  if constexpr (BIT == 5)
    xl_bit5(a, b);
  else if constexpr (BIT == 4)
    xl_bit4(a, b);
  else
    xl_bit3(a, b);
}

template <int BIT>
__device__ __forceinline__ void xl_val(double& a, double& b) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  unsigned al = (unsigned)ua, ah = (unsigned)(ua >> 32), bl = (unsigned)ub, bh = (unsigned)(ub >> 32);
  xl_word<BIT>(al, bl);
  xl_word<BIT>(ah, bh);
  a = __longlong_as_double(((unsigned long long)ah << 32) | al);
  b = __longlong_as_double(((unsigned long long)bh << 32) | bl);
}
template <int BIT>
__device__ __forceinline__ void xl_val(float& a, float& b) {
  unsigned ua = __float_as_uint(a), ub = __float_as_uint(b);
  xl_word<BIT>(ua, ub);
  a = __uint_as_float(ua);
  b = __uint_as_float(ub);
}

// Transpose lane bits 3-5 with the register index of v (8 values a lane).
template <typename R>
__device__ __forceinline__ void xl_transpose_hi(cpx<R> (&v)[8]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {  // lane bit 5 <-> register bit 2
    xl_val<5>(v[r].x, v[r + 4].x);
    xl_val<5>(v[r].y, v[r + 4].y);
  }
#pragma unroll
  for (int r : {0, 1, 4, 5}) {  // lane bit 4 <-> register bit 1
    xl_val<4>(v[r].x, v[r + 2].x);
    xl_val<4>(v[r].y, v[r + 2].y);
  }
#pragma unroll
  for (int r : {0, 2, 4, 6}) {  // lane bit 3 <-> register bit 0
    xl_val<3>(v[r].x, v[r + 1].x);
    xl_val<3>(v[r].y, v[r + 1].y);
  }
}

// v[k] *= w^k for k = 1..7, with w^k formed by complex products from w
// (w^2 = w w, w^4 = w^2 w^2, w^3, w^5, w^6, w^7 from those): one table read
// instead of seven.
template <typename R>
__device__ __forceinline__ void twiddle_powers(cpx<R> (&v)[8], cpx<R> w1) {
  const cpx<R> w2 = w1 * w1, w4 = w2 * w2;
  const cpx<R> w3 = w2 * w1, w5 = w4 * w1, w6 = w4 * w2, w7 = w4 * w3;
  v[1] = v[1] * w1;
  v[2] = v[2] * w2;
  v[3] = v[3] * w3;
  v[4] = v[4] * w4;
  v[5] = v[5] * w5;
  v[6] = v[6] * w6;
  v[7] = v[7] * w7;
}

// The transform: the contract of fft_regs_c with N = 512, one group per
// wavefront (gt = lane), tw[t] = exp(-2 pi i t / 512).  TWP: the twiddles of
// a step are formed as powers of one table entry (twiddle_powers) instead of
// seven table reads.
template <typename R, int S, bool TWP = false>
__device__ __forceinline__ void fft512_xl(cpx<R> (&v)[8], cpx<R>* buf, const cpx<R>* __restrict__ tw, int gt) {
  auto twid = [&](int t) {
    cpx<R> c = tw[t & 511];
    if (S > 0) c.y = -c.y;
    return c;
  };
  dft8<R, S>(v);  // step A: v[k2]
  if constexpr (TWP) {
    twiddle_powers(v, twid(gt));
  } else {
#pragma unroll
    for (int k = 1; k < 8; ++k) v[k] = v[k] * twid(gt * k);
  }
  xl_transpose_hi(v);  // lane bits 3-5 = k2, v[n1]
  dft8<R, S>(v);       // step B: v[k1]
  const int n0 = gt & 7, hi = gt >> 3;
  if constexpr (TWP) {
    twiddle_powers(v, twid(8 * n0));
  } else {
#pragma unroll
    for (int k = 1; k < 8; ++k) v[k] = v[k] * twid(8 * n0 * k);
  }
  // low lane bits <-> register through LDS, element n0 + 8 k1 + 64 k2
  fft_sync<true>();  // the previous use of buf is finished
  const int pw = pidx(n0 + 64 * hi);
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[pw + 9 * k] = v[k];  // pidx(e + 8k) = pidx(e) + 9k
  fft_sync<true>();
  const int pr = pidx(8 * n0 + 64 * hi);
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[pr + r];  // n0' = r: element r + 8 k1 + 64 k2
  dft8<R, S>(v);  // step C: v[k0]
  fft_sync<true>();
  const int po = pidx(hi + 8 * n0);  // k = k2 + 8 k1 + 64 k0, lane = k1 + 8 k2
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[po + 72 * k] = v[k];  // pidx(e + 64k) = pidx(e) + 72k
  fft_sync<true>();
}

// The LOGN of fft_any that selects fft512_xl (N = 512), with its twiddles
// read from the table (FFT_XL512) or formed as powers (FFT_XL512P).
constexpr int FFT_XL512 = 100, FFT_XL512P = 101;

// fft_regs, or fft_regs_c when LOGN > 0 fixes N = 2^LOGN at compile time, or
// fft512_xl for LOGN = FFT_XL512 or FFT_XL512P.
template <typename R, int S, bool WAVE, int LOGN>
__device__ __forceinline__ void fft_any(cpx<R> (&v)[8], cpx<R>* buf, int N, int logN,
                                        const cpx<R>* __restrict__ tw, int gt) {
  if constexpr (LOGN == FFT_XL512)
    fft512_xl<R, S>(v, buf, tw, gt);
  else if constexpr (LOGN == FFT_XL512P)
    fft512_xl<R, S, true>(v, buf, tw, gt);
  else if constexpr (LOGN > 0)
    fft_regs_c<R, S, WAVE, LOGN>(v, buf, tw, gt);
  else
    fft_regs<R, S, WAVE>(v, buf, N, logN, tw, gt);
}

// The n-entry twiddle table copied into LDS at dst (past the transform
// buffers), for the passes of every transform the workgroup runs: their
// twiddle reads then come at LDS latency, not the caches' (at 512x257x512
// the products pass went from 6.89 to 5.85 ms per step in FP64, 3.88 to
// 2.91 in FP32).  Every thread of the workgroup must call it.
template <typename R>
__device__ __forceinline__ const cpx<R>* lds_twiddles(cpx<R>* dst, const cpx<R>* __restrict__ tw, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = tw[t];
  __syncthreads();
  return dst;
}

// fft_regs with the wave/workgroup choice made from N: transforms of at most
// 512 points (G <= 64 lanes) synchronise within their wavefront.
template <typename R, int S>
__device__ __forceinline__ void fft_regs_auto(cpx<R> (&v)[8], cpx<R>* buf, int N, int logN,
                                              const cpx<R>* __restrict__ tw, int gt) {
  if (N <= 512)
    fft_regs<R, S, true>(v, buf, N, logN, tw, gt);
  else
    fft_regs<R, S, false>(v, buf, N, logN, tw, gt);
}

}  // namespace dgpu
