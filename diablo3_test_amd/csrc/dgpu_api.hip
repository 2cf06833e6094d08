// dgpu_api.hip -- host side of the MI355X channel substep: the handle, device
// memory, the Crank-Nicolson factor tables, the order of kernel launches for
// one rk_chan_1 call (for/channel.f90:10-917), and the extern "C" entry points
// declared in include/diablo_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/diablo_gpu.h"
#include "dgpu_adjoint.h"
#include "dgpu_init.h"
#include "dgpu_les.h"
#include "dgpu_les_adj.h"
#include "dgpu_zgen.h"

using namespace dgpu;

namespace {

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIPCHECK(call)                             \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) throw HipError{e_, #call}; \
  } while (0)

struct ApiError {
  int code;
  std::string msg;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    n = count;
    if (count) HIPCHECK(hipMalloc(&p, sizeof(T) * count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Tuning knobs.  Each handle reads them once, at dgpu_create, from the
// environment (DGPU_SEG_TARGET, DGPU_XPASS_THREADS, DGPU_XPROD_THREADS,
// DGPU_LES_COLS, DGPU_THOMAS_COLS, DGPU_SPECTRAL_CN, DGPU_BATCH_MB,
// DGPU_FFT_CONST), with the
// defaults below when a variable is unset, and keeps its own copy: creating a
// handle never changes how another one runs.
struct Knobs {
  // Rows per segment of the CN solver: one of the instantiated lengths (the
  // rows live in registers), with at most 16 segments (1024 threads per
  // block).  The target length trades registers per thread against segments
  // per column.  At 512x257x512 20 rows (13 segments) took the spectral
  // solves from 2.47 to 2.34 ms per step against 24 (11 segments): the kernel
  // runs 4 waves per SIMD, so a CU holds one block, and 13 waves fill more of
  // its 16 slots; 17 rows (16 segments, 97 VGPRs) fill all 16 and take them
  // from 2.24 to 2.17 (mixed 2.12 to 2.04).
  int seg_target = 17;
  // Threads per workgroup of the x passes and of the products pass (FP64,
  // FP32).  The FP64 products pass holds 228 VGPRs, two waves per SIMD: as
  // two 256-thread workgroups per CU (each with 4 of the 8 rows of its tile,
  // in two rounds) it measured 4.79 -> 4.71 ms per step at 512x257x512
  // against one 512-thread group; the FP32 pass, whose 16-row tiles would
  // then be written as 32-byte runs, keeps 512 (2.30 ms, 6.67 with 256).
  int xpass_threads = 512;
  int xprod_threads = 256;
  int xprod_threads32 = 512;
  // Columns per lane of the LES CN solve.
  int les_cols = 2;
  // Columns per lane of the FP32-storage CN solve (1 or 2).  Measured at
  // 512x257x512: 2 columns take the mixed form's 12 solves per step from 5.19
  // to 4.50 ms; the FP64 solves are 2% slower with 2 (5.21 -> 5.33 ms).
  int thomas_cols = 2;
  // CN solves of the non-LES substep in spectral space (0 = the
  // physical-space solves between the transforms, as the reference).
  int spectral_cn = 1;
  // Plane batches of a 2-D transform: the x and z passes run batch by batch
  // so that a batch's x-spectra (the intermediate between the passes) are
  // read back while still in the last-level cache.  Budget in MB.  Off by
  // default: at 512x257x512 every batch size measured slower than one launch
  // per pass over all planes (state to physical space, FP64 ms per step:
  // 64 MB a field 2.60 + 1.67 for the x and z inverses, 192 MB 2.22 + 1.32,
  // against 2.10 + 1.24 unbatched; batches of all four fields at once sized
  // for the 256 MB last-level cache, 2.17 + 1.30 at 320 MB).
  int batch_mb = 0;
  // Transform passes over 512 points with the length fixed at compile time
  // (fft_regs_c): bit 1 the FP64 passes, bit 2 the FP32 RHS pass (measured
  // at 512x257x512: 3.64 -> 3.58 ms per step); 0 = the run-time length
  // everywhere.
  int fft_const = 3;
  // Products pass with each input phase's loads issued behind the previous
  // pair's products, before its transform: bit 1 FP64, bit 2 FP32 (the FP32
  // pass with its 512-point cross-lane transform; 0 = at the phase start).
  int xprod_pf = 3;
  // 512-point passes that run fft512_xl (the cross-lane exchange), a bit per
  // pass (XL_* below): the products, x and z inverse and x and z forward
  // passes in FP64 and FP32 and the adjoint's x pass (1 + 4 + 8 + 16 + 32 +
  // 64 + 128).  Measured at 512x257x512: DNS FP64 16.72 -> 16.32 ms per step,
  // mixed 11.74 -> 11.50, LES 60.6 -> 59.4, adjoint 71.5 -> 70.9.  The RHS
  // pass keeps the Stockham form: with fft512_xl its 137 VGPRs cost a wave per
  // SIMD (4 -> 3) and it went 4.20 -> 5.66 ms.
  int fft_xl = 253;
  // The FP64 512-point passes whose fft512_xl forms its twiddles as powers of
  // one table entry (same bits): the x passes -- products, x inverse, x
  // forward, the adjoint's x pass -- and the z inverses (1 + 4 + 32 + 128 +
  // 8).  Measured at 512x257x512 (FP64 ms per step): products 4.88 -> 4.66, x
  // inverse 2.27 -> 2.10.  Except in the products pass, the power forms read
  // their one table entry a step from the caches and keep no table in LDS,
  // which fits a fourth workgroup per CU in the z passes: the DNS z inverse
  // stays at 1.24, the adjoint's z inverses go 5.78 -> 5.14 and the LES ones
  // 3.21 -> 2.78; the z forward passes are split (adjoint 1.95 -> 1.78, LES
  // 4.04 -> 4.20) and keep the table.  The FP32 passes are unchanged either
  // way.
  int fft_twp = 173;
};

Knobs knobs_from_env() {
  Knobs k;
  auto rd = [](const char* name, int& v) {
    if (const char* e = std::getenv(name)) v = std::atoi(e);
  };
  rd("DGPU_SEG_TARGET", k.seg_target);
  rd("DGPU_XPASS_THREADS", k.xpass_threads);
  rd("DGPU_XPROD_THREADS", k.xprod_threads);
  rd("DGPU_XPROD_THREADS32", k.xprod_threads32);
  rd("DGPU_LES_COLS", k.les_cols);
  rd("DGPU_THOMAS_COLS", k.thomas_cols);
  rd("DGPU_SPECTRAL_CN", k.spectral_cn);
  rd("DGPU_BATCH_MB", k.batch_mb);
  rd("DGPU_FFT_CONST", k.fft_const);
  rd("DGPU_XPROD_PF", k.xprod_pf);
  rd("DGPU_FFT_XL", k.fft_xl);
  rd("DGPU_FFT_TWP", k.fft_twp);
  k.spectral_cn = k.spectral_cn != 0;
  return k;
}

int seg_len_for(int n, int seg_target) {
  static const int lens[] = {8, 16, 17, 20, 24, 28, 32, 40, 48, 56};  // the launch switches' cases
  int S = (n + seg_target - 1) / seg_target;
  if (S > 16) S = 16;
  const int L0 = (n + S - 1) / S;
  for (int L : lens)
    if (L >= L0) return (n + L - 1) / L > 16 ? -1 : L;
  return -1;
}

}  // namespace

struct dgpu_handle {
  dgpu_params p;
  int precision;
  Geo g;
  int nth;
  double nu, pi;
  std::vector<double> gy, gyf, dy, dyf, kx, kz;
  int jstart, jend, jstart_th[DGPU_MAX_TH], jend_th[DGPU_MAX_TH];
  double h_bar[3], beta_bar[3], zeta_bar[3];
  double w_transient;
  hipStream_t stream = nullptr;
  std::string err;
  Knobs kn;  // tuning knobs, read at dgpu_create

  // device state (layouts described in dgpu_kernels.h)
  DevBuf<double2> cu[3], cp, cth[DGPU_MAX_TH], cf[3], cfth[DGPU_MAX_TH];
  DevBuf<double> u[3], th[DGPU_MAX_TH], r[3], rth[DGPU_MAX_TH];
  DevBuf<double> nu_t, kappa_t[DGPU_MAX_TH];
  std::vector<DevBuf<double2>> tpool;
  // Nz not a power of two: the z passes run the general transforms of
  // dgpu_zgen.h with this plan, and zpool holds the z spectra of the product
  // slots (FP64, plane pitch sp) that the RHS pass then combines
  bool zgen = false;
  ZPlan zplan{};
  std::vector<DevBuf<double2>> zpool;
  DevBuf<double> sb;
  DevBuf<double2> sg;
  // projection: 1/b_j of every mode (rows 1..Ny) and the row coefficients
  DevBuf<double> rd_ib;
  DevBuf<PoisRow> d_prow;
  DevBuf<double2> twx, twz;
  DevBuf<float2> twx32, twz32;  // FP32 twiddles of the mixed form
  DevBuf<double> d_kx, d_kz, d_dy, d_dyf, d_gyf;
  // Front thermal wind: (gyf(j) - Ly/2) delta/Ro_inv on planes 0..Ny+1, then
  // the same at gy(j) (allocated only for flavor == 'Front' with Ro_inv != 0)
  DevBuf<double> d_front;
  // per-block flags of the non-finite scan (dgpu_check_state)
  DevBuf<int> d_flags;
  // address of the static scratch array on this handle's device
  void* scratch = nullptr;
  // ordering events of the copies on a caller's stream (dgpu_*_field_device)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // poisson_p_chan: the spectral right-hand side (allocated on first use)
  DevBuf<double2> pois;
  // variable_dt (dgpu_set_courant): settings, CFL*dy(j), Front offsets,
  // per-workgroup minima
  bool var_dt = false;
  dgpu_courant_params cour{};
  int cfl_nset = 1;
  DevBuf<double> cfl_cdy, cfl_offx, cfl_offz, cfl_part;
  // Crank-Nicolson factor tables: [field][rk] rows, field 0 u1, 1 u2, 2 u3, 3+n th_n
  DevBuf<SegRow> rows;
  // LES (allocated when use_les): velocity/scalar gradients, model outputs,
  // their spectra, the subgrid forcing, the wall planes of the pre-BC state, CN scratch
  DevBuf<TriRow> trirows;              // raw CN rows [field][rk][j]
  DevBuf<double> grad[9];
  DevBuf<double> dth[3][DGPU_MAX_TH];
  DevBuf<double> tau[7];               // t11 t33 t13 t12 t23 c1 c3
  DevBuf<double> thf[2][DGPU_MAX_TH];
  DevBuf<double2> ctau[7];
  DevBuf<double2> cthf[2][DGPU_MAX_TH];
  DevBuf<double2> lesf[3], lesfth[DGPU_MAX_TH];
  DevBuf<double2> wpre[3 + DGPU_MAX_TH];  // planes 0, 1, Ny, Ny+1 before the LES wall values
  DevBuf<double> sbl, sgl;              // b', g' of the per-column CN solve
  // adjoint (allocated by the first adjoint call): the adjoint state, its
  // scratch, the transposed CN factor tables, and forward-state checkpoints
  bool adj_ready = false;
  DevBuf<double2> a_cu[3], a_cp, a_cth[DGPU_MAX_TH], a_cf[3], a_cfth[DGPU_MAX_TH];
  DevBuf<double2> a_cr[3 + DGPU_MAX_TH], a_fb[3 + DGPU_MAX_TH];
  DevBuf<double> a_rb[3 + DGPU_MAX_TH], a_ub[3 + DGPU_MAX_TH];
  DevBuf<double> a_pA[4 + 2 * DGPU_MAX_TH], a_pB[4 + 2 * DGPU_MAX_TH];
  DevBuf<SegRow> rowsT;
  DevBuf<double> d_w;  // Ny+2 weights of dgpu_adj_add_mean
  // adjoint of the LES terms (allocated by the first LES adjoint call): the
  // forward recomputation's captures (post-BC state, CN solutions), the saved
  // forward state, the adjoint of the post-BC state, and the adjoints of
  // nu_t, kappa_t, the gradients and the scalar gradients
  bool les_cap = false;
  bool les_adj_ready = false;
  DevBuf<double2> cap_cub[3 + DGPU_MAX_TH];
  DevBuf<double> cap_x[3 + DGPU_MAX_TH];
  std::vector<DevBuf<double2>> fsave;
  DevBuf<double2> a_cub[3 + DGPU_MAX_TH];
  DevBuf<double> a_nt, a_kt[DGPU_MAX_TH], a_G[9], a_T[3][DGPU_MAX_TH];
  struct Snap {
    std::vector<DevBuf<double2>> f;
    size_t nsaved = 0;
    double w_transient = 0.0;
  };
  std::vector<Snap> snaps;
  int seg_L = 0, seg_S = 0;
  int nfields = 0;
  size_t lds_limit = 65536;
  // optional per-kernel-class timing (dgpu_timing_enable)
  bool timing = false;
  struct Mark {
    int cls;
    double bytes, tbytes;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  double t_ms[DGPU_NCLASS] = {0};
  double t_bytes[DGPU_NCLASS] = {0};
  double t_tbytes[DGPU_NCLASS] = {0};
  long long t_count[DGPU_NCLASS] = {0};
};

namespace {

// ------------------------------------------------------------ geometry ----
void setup_grid(dgpu_handle* h, const double* gy_in) {
  const dgpu_params& p = h->p;
  const int Ny = p.ny, n2 = Ny + 2;
  h->gy.assign(n2, 0.0);
  h->gyf.assign(n2, 0.0);
  h->dy.assign(n2, 0.0);
  h->dyf.assign(n2, 0.0);
  // gy(1..Ny+1) from grid.h5 (domain.f90:429-483); metrics as domain.f90:230-254, 501-527
  for (int j = 1; j <= Ny + 1; ++j) h->gy[j] = gy_in[j - 1];
  for (int j = 1; j <= Ny; ++j) h->gyf[j] = 0.5 * (h->gy[j] + h->gy[j + 1]);
  h->gyf[0] = 2.0 * h->gyf[1] - h->gyf[2];
  h->gyf[Ny + 1] = 2.0 * h->gyf[Ny] - h->gyf[Ny - 1];
  for (int j = 1; j <= Ny + 1; ++j) h->dy[j] = h->gyf[j] - h->gyf[j - 1];
  for (int j = 1; j <= Ny; ++j) h->dyf[j] = h->gy[j + 1] - h->gy[j];
  h->dyf[0] = h->dyf[1];
  h->dyf[Ny + 1] = h->dyf[Ny];
  // wavenumbers (fft.f90:152, 213-230); pi = 4*atan(1.0) in default real kind
  h->pi = (double)(4.0f * std::atan(1.0f));
  const int nkz = p.nz / 3;
  h->kx.assign(h->g.nq, 0.0);
  for (int q = 0; q < h->g.nq; ++q) h->kx[q] = (double)q * (2.0 * h->pi) / p.lx;
  h->kz.assign(2 * nkz + 1, 0.0);
  for (int k = 0; k <= nkz; ++k) h->kz[k] = (double)k * (2.0 * h->pi) / p.lz;
  for (int k = 1; k <= nkz; ++k) h->kz[2 * nkz + 1 - k] = -(double)k * (2.0 * h->pi) / p.lz;
  // j ranges at one y-rank (domain.f90:284-303): jend = Ny whatever the top BC
  h->jend = Ny;
  h->jstart = (p.u_bc_ymin == 1) ? 1 : 2;
  for (int n = 0; n < p.n_th; ++n) {
    h->jend_th[n] = Ny;
    h->jstart_th[n] = (p.th_bc_ymin[n] == 1) ? 1 : 2;
  }
  // RK3 (flow.f90:78-86)
  h->h_bar[0] = p.delta_t * (8.0 / 15.0);
  h->h_bar[1] = p.delta_t * (2.0 / 15.0);
  h->h_bar[2] = p.delta_t * (5.0 / 15.0);
  h->beta_bar[0] = 1.0;
  h->beta_bar[1] = 25.0 / 8.0;
  h->beta_bar[2] = 9.0 / 4.0;
  h->zeta_bar[0] = 0.0;
  h->zeta_bar[1] = -17.0 / 8.0;
  h->zeta_bar[2] = -5.0 / 4.0;
  h->w_transient = p.w_bc_ymax_c1;
}

// The matrix of one real CN solve (rows j = 0..Ny+1) before factorisation.
struct Tri {
  std::vector<double> a, b, c, vb, wp, wm;
  std::vector<int> kind;
};

// channel.f90:656-707 (th), :716-765 (u2), :771-822 (u1), :825-877 (u3) without
// LES, then the BC rows of apply_BC_{1,2,3,th}_{lower,upper} (:1545-1876).
Tri build_tri(const dgpu_handle* h, int field, int rs) {
  const dgpu_params& p = h->p;
  const int Ny = p.ny, n = Ny + 2;
  Tri t;
  t.a.assign(n, 0.0);
  t.b.assign(n, 1.0);
  t.c.assign(n, 0.0);
  t.vb.assign(n, 0.0);
  t.wp.assign(n, 0.0);
  t.wm.assign(n, 0.0);
  t.kind.assign(n, 0);
  const double temp1 = 0.5 * h->nu * h->h_bar[rs];
  const auto& dy = h->dy;
  const auto& dyf = h->dyf;
  int ja, jb;
  double ce;
  if (field == 1) {
    ja = 2;
    jb = Ny;
    ce = temp1 * p.nu_v_scale;
    for (int j = ja; j <= jb; ++j) {
      t.a[j] = -ce / (dyf[j - 1] * dy[j]);
      t.b[j] = 1.0 + ce / (dyf[j] * dy[j]) + ce / (dyf[j - 1] * dy[j]);
      t.c[j] = -ce / (dyf[j] * dy[j]);
      t.wp[j] = ce / (dyf[j] * dy[j]);
      t.wm[j] = ce / (dyf[j - 1] * dy[j]);
      t.kind[j] = 1;
    }
  } else {
    if (field >= 3) {
      const int nn = field - 3;
      ja = h->jstart_th[nn];
      jb = h->jend_th[nn];
      ce = temp1 / p.pr[nn] * p.nu_v_scale;
    } else {
      ja = h->jstart;
      jb = h->jend;
      ce = temp1 * p.nu_v_scale;
    }
    for (int j = ja; j <= jb; ++j) {
      t.a[j] = -ce / (dy[j] * dyf[j]);
      t.b[j] = 1.0 + ce / (dy[j + 1] * dyf[j]) + ce / (dy[j] * dyf[j]);
      t.c[j] = -ce / (dy[j + 1] * dyf[j]);
      t.wp[j] = ce / (dy[j + 1] * dyf[j]);
      t.wm[j] = ce / (dy[j] * dyf[j]);
      t.kind[j] = 1;
    }
  }
  auto row = [&](int j, double l, double d, double u, double v, int kind = 0) {
    t.a[j] = l;
    t.b[j] = d;
    t.c[j] = u;
    t.vb[j] = v;
    t.kind[j] = kind;
  };
  if (field == 1) {
    if (p.v_bc_ymin == 0) {
      row(1, 0, 1, 0, p.v_bc_ymin_c1);
      row(2, 0, 1, 0, p.v_bc_ymin_c1);
    } else if (p.v_bc_ymin == 1) {
      row(1, 0, -1, 1, dyf[1] * p.v_bc_ymin_c1);
    }
    row(0, 0, 1, 0, 0);
    if (p.v_bc_ymax == 0) {
      row(Ny + 1, 0, 1, 0, p.v_bc_ymax_c1);
      row(Ny, 0, 1, 0, p.v_bc_ymax_c1);
    } else if (p.v_bc_ymax == 1) {
      row(Ny + 1, -1, 1, 0, dyf[Ny] * p.v_bc_ymax_c1);
    }
  } else {
    int bcl, bcu;
    double c1l, c1u;
    bool transient = false;
    if (field == 0) {
      bcl = p.u_bc_ymin; bcu = p.u_bc_ymax; c1l = p.u_bc_ymin_c1; c1u = p.u_bc_ymax_c1;
    } else if (field == 2) {
      bcl = p.w_bc_ymin; bcu = p.w_bc_ymax; c1l = p.w_bc_ymin_c1; c1u = p.w_bc_ymax_c1;
      transient = (bcu != 0);  // uses w_BC_Ymax_c1_transient (channel.f90:1775-1780)
    } else {
      const int nn = field - 3;
      bcl = p.th_bc_ymin[nn]; bcu = p.th_bc_ymax[nn];
      c1l = p.th_bc_ymin_c1[nn]; c1u = p.th_bc_ymax_c1[nn];
    }
    row(0, 0, 1, 0, 0);
    if (bcl == 0)
      row(1, 0, 1, 0, c1l);
    else
      row(1, 0, -1, 1, dy[2] * c1l);
    if (bcu == 0) {
      row(Ny + 1, 0, 1, 0, 0);
      row(Ny, 0, 1, 0, c1u);
    } else {
      if (transient)
        row(Ny, -1, 1, 0, dy[Ny], 2);  // vb scaled by the current transient value
      else
        row(Ny, -1, 1, 0, dy[Ny] * c1u);
      row(Ny + 1, 0, 1, 0, 0);
    }
  }
  return t;
}

// LU factors of the Thomas sweep and the segment carry products.
void factor_rows(const Tri& t, int L, SegRow* out) {
  const int n = (int)t.a.size();
  std::vector<double> bp(n), m(n, 0.0);
  bp[0] = t.b[0];
  for (int j = 1; j < n; ++j) {
    m[j] = -t.a[j] / bp[j - 1];
    bp[j] = t.b[j] + m[j] * t.c[j - 1];
  }
  for (int j = 0; j < n; ++j) {
    SegRow& r = out[j];
    r.m = m[j];
    r.e = 1.0 / bp[j];
    r.f = -t.c[j] / bp[j];
    r.wp = t.wp[j];
    r.wm = t.wm[j];
    r.kind = t.kind[j];
    r.vb = t.vb[j];
    r.ra = t.a[j];
    r.rb = t.b[j];
    r.rc = t.c[j];
  }
  for (int s0 = 0; s0 < n; s0 += L) {
    const int s1 = std::min(s0 + L, n) - 1;
    double prod = 1.0;
    for (int j = s0; j <= s1; ++j) {
      prod *= out[j].m;
      out[j].pm = prod;
    }
    prod = 1.0;
    for (int j = s1; j >= s0; --j) {
      prod *= out[j].f;
      out[j].qf = prod;
    }
  }
}

// Raw CN rows with the LES weights of channel.f90:676-689 (th), 736-749 (u2),
// 792-805 (u1), 848-861 (u3): temp2 = h/2 times the face metrics.
void raw_rows(const dgpu_handle* h, const Tri& t, int field, int rs, TriRow* out) {
  const int n = (int)t.a.size();
  const double temp2 = 0.5 * h->h_bar[rs];
  const auto& dy = h->dy;
  const auto& dyf = h->dyf;
  for (int j = 0; j < n; ++j) {
    TriRow& r = out[j];
    r.a = t.a[j];
    r.b = t.b[j];
    r.c = t.c[j];
    r.vb = t.vb[j];
    r.wp = t.wp[j];
    r.wm = t.wm[j];
    r.kind = t.kind[j];
    r.vtype = (field == 1);
    r.lp = r.lm = 0.0;
    if (t.kind[j] == 1) {
      if (field == 1) {
        r.lp = temp2 / (dyf[j] * dy[j]);
        r.lm = temp2 / (dyf[j - 1] * dy[j]);
      } else {
        r.lp = temp2 / (dy[j + 1] * dyf[j]);
        r.lm = temp2 / (dy[j] * dyf[j]);
      }
    }
  }
}

void twiddles(int n, std::vector<double2>& w) {
  w.resize(n);
  for (int t = 0; t < n; ++t) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
    w[t] = make_double2((double)cosl(a), (double)sinl(a));
  }
}

template <typename T>
void upload(DevBuf<T>& b, const std::vector<T>& v) {
  if (!b.p || b.n != v.size()) {
    b.release();
    b.alloc(v.size());
  }
  HIPCHECK(hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
}

// The Crank-Nicolson tables of every field and RK substep for the current
// h_bar: the segmented factorisation, and with use_LES the raw rows.
void build_cn_tables(dgpu_handle* h) {
  const int n = h->g.ny2;
  std::vector<SegRow> rows((size_t)h->nfields * 3 * n);
  for (int f = 0; f < h->nfields; ++f)
    for (int rs = 0; rs < 3; ++rs) {
      Tri t = build_tri(h, f, rs);
      factor_rows(t, h->seg_L, rows.data() + ((size_t)f * 3 + rs) * n);
    }
  upload(h->rows, rows);
  if (h->p.use_les) {
    std::vector<TriRow> tr((size_t)h->nfields * 3 * n);
    for (int f = 0; f < h->nfields; ++f)
      for (int rs = 0; rs < 3; ++rs) {
        Tri t = build_tri(h, f, rs);
        raw_rows(h, t, f, rs, tr.data() + ((size_t)f * 3 + rs) * n);
      }
    upload(h->trirows, tr);
  }
}

// ------------------------------------------------------------- timing ----
hipEvent_t take_event(dgpu_handle* h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHECK(hipEventCreate(&e));
  return e;
}

// Brackets one kernel launch with events when timing is on.  `bytes` is the
// launch's algorithmic HBM traffic (compulsory reads + writes).
// tbytes: the part of `bytes` that is x-spectra traffic (the intermediate
// of a two-pass 2-D transform, written by one pass and read by the next),
// which a one-pass transform would not move.
struct Timed {
  dgpu_handle* h;
  int cls;
  double bytes, tbytes;
  hipEvent_t a = nullptr;
  Timed(dgpu_handle* h_, int cls_, double bytes_, double tbytes_ = 0.0)
      : h(h_), cls(cls_), bytes(bytes_), tbytes(tbytes_) {
    if (h->timing) {
      a = take_event(h);
      HIPCHECK(hipEventRecord(a, h->stream));
    }
  }
  ~Timed() {
    // never throws: a failed launch has already raised, and its events are
    // simply returned to the pool
    if (h->timing && a) {
      hipEvent_t b = nullptr;
      if (!h->event_pool.empty()) {
        b = h->event_pool.back();
        h->event_pool.pop_back();
      } else if (hipEventCreate(&b) != hipSuccess) {
        b = nullptr;
      }
      if (b && hipEventRecord(b, h->stream) == hipSuccess) {
        h->marks.push_back({cls, bytes, tbytes, a, b});
      } else {
        h->event_pool.push_back(a);
        if (b) h->event_pool.push_back(b);
      }
    }
  }
};

void collect_timing(dgpu_handle* h) {
  if (h->marks.empty()) return;
  HIPCHECK(hipStreamSynchronize(h->stream));
  for (auto& m : h->marks) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, m.a, m.b));
    h->t_ms[m.cls] += ms;
    h->t_bytes[m.cls] += m.bytes;
    h->t_tbytes[m.cls] += m.tbytes;
    h->t_count[m.cls] += 1;
    h->event_pool.push_back(m.a);
    h->event_pool.push_back(m.b);
  }
  h->marks.clear();
}

// ------------------------------------------------------- launch helpers ----
// R is the transform precision: double for the FP64 path, float for the mixed
// form.  Physical fields and x-spectra are stored in R (the buffers are sized
// for FP64 and a float view uses their first half); spectral fields are FP64.
template <typename R>
R* phys(const DevBuf<double>& b) {
  return reinterpret_cast<R*>(b.p);
}
template <typename R>
cpx<R>* xspec(const DevBuf<double2>& b) {
  return reinterpret_cast<cpx<R>*>(b.p);
}
template <typename R>
const cpx<R>* twx_of(const dgpu_handle* h);
template <>
const cpx<double>* twx_of<double>(const dgpu_handle* h) {
  return reinterpret_cast<const cpx<double>*>(h->twx.p);
}
template <>
const cpx<float>* twx_of<float>(const dgpu_handle* h) {
  return reinterpret_cast<const cpx<float>*>(h->twx32.p);
}
template <typename R>
const cpx<R>* twz_of(const dgpu_handle* h);
template <>
const cpx<double>* twz_of<double>(const dgpu_handle* h) {
  return reinterpret_cast<const cpx<double>*>(h->twz.p);
}
template <>
const cpx<float>* twz_of<float>(const dgpu_handle* h) {
  return reinterpret_cast<const cpx<float>*>(h->twz32.p);
}

// Transform groups per workgroup for the x passes: a divisor of `jobs` (the
// transforms one KT-row tile needs), at most max_threads threads, within the
// LDS limit.  All of a tile's transforms in one round when they fit.
template <typename R>
int x_ngroups(const dgpu_handle* h, int jobs, int max_threads = 512, size_t extra = 0) {
  const Geo& g = h->g;
  const int G = g.nx / 8;
  int ng = jobs;
  // the group buffers and the twiddle table (lds_twiddles), plus extra bytes
  auto lds = [&](int n) { return ((size_t)n * x_group_stride(g.nx) + g.nx) * sizeof(cpx<R>) + extra; };
  while (ng > 1 && (ng * G > max_threads || lds(ng) > h->lds_limit)) ng /= 2;
  if (ng * G > max_threads || lds(ng) > h->lds_limit)
    throw ApiError{DGPU_EUNSUP, "Nx too large for the x-transform pass"};
  return ng;
}

// Whether a 512-point pass in precision R takes fft512_xl (knob fft_xl: a
// bit per pass, XL_*, plus XL_FP32 for the FP32 passes).
enum { XL_PROD = 1, XL_RHS = 2, XL_XINV = 4, XL_ZINV = 8, XL_FP32 = 16, XL_XFWD = 32, XL_ZFWD = 64, XL_ADJ = 128 };
template <typename R>
bool use_xl(const dgpu_handle* h, int pass) {
  return (h->kn.fft_xl & pass) && (sizeof(R) == 8 || (h->kn.fft_xl & XL_FP32));
}
// Whether that pass forms its twiddles as powers (knob fft_twp, FP64 only).
template <typename R>
bool xl_powers(const dgpu_handle* h, int pass) {
  return sizeof(R) == 8 && (h->kn.fft_twp & pass);
}

template <typename R>
size_t x_lds(const dgpu_handle* h, int ng, bool table = true) {
  return ((size_t)ng * x_group_stride(h->g.nx) + (table ? h->g.nx : 0)) * sizeof(cpx<R>);
}

template <typename R>
int z_ngroups(const dgpu_handle* h) {
  const Geo& g = h->g;
  const int G = g.nz / 8;
  if (G > 512) throw ApiError{DGPU_EUNSUP, "Nz too large for the z-transform pass"};
  int ng = 256 / G;
  if (ng < 1) ng = 1;
  auto lds = [&](int n) { return ((size_t)n * fft_lds_len(g.nz) + g.nz) * sizeof(cpx<R>); };
  while (ng > 1 && lds(ng) > h->lds_limit) --ng;
  if (lds(ng) > h->lds_limit) throw ApiError{DGPU_EUNSUP, "Nz too large for the z-transform pass"};
  return ng;
}

template <typename R>
size_t z_lds(const dgpu_handle* h, int ng, bool table = true) {
  return ((size_t)ng * fft_lds_len(h->g.nz) + (table ? h->g.nz : 0)) * sizeof(cpx<R>);
}

template <typename R>
void launch_xfwd(dgpu_handle* h, const R* u, cpx<R>* T, int j0, int j1, const FieldBatch* fb = nullptr) {
  if (j1 < j0) return;
  const Geo& g = h->g;
  constexpr int KT = XTile<R>::KT;
  const int ng = x_ngroups<R>(h, KT / 2, h->kn.xpass_threads);
  const int nf = fb ? fb->nf : 1;
  dim3 grid((g.nz + KT - 1) / KT, j1 - j0 + 1, nf), block(ng * (g.nx / 8));
  Timed tm(h, DGPU_K_XFWD, (sizeof(R) * g.pp + 2.0 * sizeof(R) * g.tp) * (j1 - j0 + 1) * nf,
           2.0 * sizeof(R) * g.tp * (j1 - j0 + 1) * nf);
  auto kern = g.nx <= 512 ? k_xfwd<R, true> : k_xfwd<R, false>;
  if (g.nx == 512 && use_xl<R>(h, XL_XFWD))
    kern = xl_powers<R>(h, XL_XFWD) ? k_xfwd<R, true, false, FFT_XL512P> : k_xfwd<R, true, false, FFT_XL512>;
  if (g.nz % KT) kern = g.nx <= 512 ? k_xfwd<R, true, true> : k_xfwd<R, false, true>;  // ragged last tile
  // the power-twiddle form keeps no table in LDS
  const size_t lds = x_lds<R>(h, ng, kern != k_xfwd<R, true, false, FFT_XL512P>);
  hipLaunchKernelGGL(kern, grid, block, lds, h->stream, u, T, g, twx_of<R>(h), j0, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

template <typename R>
void launch_xinv(dgpu_handle* h, const cpx<R>* T, R* u, int j0, int j1, const double* dkx = nullptr,
                 const FieldBatch* fb = nullptr) {
  if (j1 < j0) return;
  const Geo& g = h->g;
  constexpr int KT = XTile<R>::KT;
  const int ng = x_ngroups<R>(h, KT / 2, h->kn.xpass_threads);
  const size_t lds = x_lds<R>(h, ng, XInvTable<R>::on);
  const int nf = fb ? fb->nf : 1;
  dim3 grid((g.nz + KT - 1) / KT, j1 - j0 + 1, nf), block(ng * (g.nx / 8));
  Timed tm(h, DGPU_K_XINV, (sizeof(R) * g.pp + 2.0 * sizeof(R) * g.tp) * (j1 - j0 + 1) * nf,
           2.0 * sizeof(R) * g.tp * (j1 - j0 + 1) * nf);
  auto kern = g.nx <= 512 ? k_xinv<R, true> : k_xinv<R, false>;
  if constexpr (sizeof(R) == 8)
    if (g.nx == 512 && (h->kn.fft_const & 1)) kern = k_xinv<R, true, 9>;
  if (g.nx == 512 && use_xl<R>(h, XL_XINV))
    kern = xl_powers<R>(h, XL_XINV) ? k_xinv<R, true, FFT_XL512P> : k_xinv<R, true, FFT_XL512>;
  if (g.nz % KT) kern = g.nx <= 512 ? k_xinv<R, true, 0, true> : k_xinv<R, false, 0, true>;  // ragged last tile
  hipLaunchKernelGGL(kern, grid, block, lds, h->stream, T, u, g, twx_of<R>(h), j0, dkx, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

// --- z transforms of any length (Nz not a power of two)
// Radices of n for the general z transforms: 4s first, then 2, then the odd
// primes in increasing order.
ZPlan make_zplan(int n) {
  ZPlan pl{};
  auto push = [&](int r) {
    if (pl.np >= ZGEN_MAXPASS) throw ApiError{DGPU_EINVAL, "nz has too many factors"};
    pl.rad[pl.np++] = r;
  };
  while (n % 4 == 0) { push(4); n /= 4; }
  while (n % 2 == 0) { push(2); n /= 2; }
  for (int f = 3; n > 1; f += 2)
    while (n % f == 0) { push(f); n /= f; }
  return pl;
}

// Threads per column and columns per workgroup of the general z passes.
template <typename R>
void zgen_shape(const dgpu_handle* h, int& tc, int& cb, size_t& lds) {
  const int n = h->g.nz;
  tc = n <= 512 ? 64 : 256;
  cb = 256 / tc;
  const size_t per = 2 * (size_t)n * sizeof(cpx<R>);
  while (cb > 1 && cb * per > h->lds_limit) --cb;
  if (cb * per > h->lds_limit) throw ApiError{DGPU_EUNSUP, "Nz too large for the z-transform pass"};
  lds = cb * per;
}

template <typename R>
void zgen_fwd(dgpu_handle* h, const cpx<R>* T, double2* C, int j0, int j1, bool accum, double s0, double s1,
              const FieldBatch* fb) {
  const Geo& g = h->g;
  int tc, cb;
  size_t lds;
  zgen_shape<R>(h, tc, cb, lds);
  const int nf = fb ? fb->nf : 1;
  dim3 grid((g.nq + cb - 1) / cb, j1 - j0 + 1, nf), block(tc * cb);
  hipLaunchKernelGGL(k_zgen_fwd<R>, grid, block, lds, h->stream, T, reinterpret_cast<cpx<double>*>(C), g, h->zplan,
                     twz_of<R>(h), j0, accum ? 1 : 0, (R)s0, (R)s1, tc, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

template <typename R>
void zgen_inv(dgpu_handle* h, const double2* C, cpx<R>* T, int j0, int j1, int op, double s0, double s1,
              const FieldBatch* fb) {
  const Geo& g = h->g;
  int tc, cb;
  size_t lds;
  zgen_shape<R>(h, tc, cb, lds);
  const int nf = fb ? fb->nf : 1;
  dim3 grid((g.nq + cb - 1) / cb, j1 - j0 + 1, nf), block(tc * cb);
  hipLaunchKernelGGL(k_zgen_inv<R>, grid, block, lds, h->stream, reinterpret_cast<const cpx<double>*>(C), T, g,
                     h->zplan, twz_of<R>(h), j0, op, h->d_kx.p, h->d_kz.p, h->d_dy.p, h->d_dyf.p, s0, s1, tc,
                     fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

template <typename R>
void launch_zfwd(dgpu_handle* h, const cpx<R>* T, double2* C, int j0, int j1, bool accum, double s0 = 1.0,
                 double s1 = 1.0, const FieldBatch* fb = nullptr) {
  if (j1 < j0) return;
  const Geo& g = h->g;
  const int nf = fb ? fb->nf : 1;
  Timed tm(h, DGPU_K_ZFWD, (2.0 * sizeof(R) * g.tp + (accum ? 32.0 : 16.0) * g.sp) * (j1 - j0 + 1) * nf,
           2.0 * sizeof(R) * g.tp * (j1 - j0 + 1) * nf);
  if (h->zgen) {
    zgen_fwd<R>(h, T, C, j0, j1, accum, s0, s1, fb);
    return;
  }
  const int ng = z_ngroups<R>(h);
  dim3 grid((g.nq + ng - 1) / ng, j1 - j0 + 1, nf), block(ng * (g.nz / 8));
  auto kern = g.nz <= 512 ? k_zfwd<R, true> : k_zfwd<R, false>;
  const bool twp = g.nz == 512 && use_xl<R>(h, XL_ZFWD) && xl_powers<R>(h, XL_ZFWD);
  if (g.nz == 512 && use_xl<R>(h, XL_ZFWD)) kern = twp ? k_zfwd<R, true, FFT_XL512P> : k_zfwd<R, true, FFT_XL512>;
  const size_t lds = z_lds<R>(h, ng, !twp);
  hipLaunchKernelGGL(kern, grid, block, lds, h->stream, T, reinterpret_cast<cpx<double>*>(C), g, twz_of<R>(h), j0,
                     accum ? 1 : 0, (R)s0, (R)s1, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

template <typename R>
void launch_zinv(dgpu_handle* h, const double2* C, cpx<R>* T, int j0, int j1, double s0 = 1.0,
                 double s1 = 1.0, const FieldBatch* fb = nullptr) {
  if (j1 < j0) return;
  const Geo& g = h->g;
  const int nf = fb ? fb->nf : 1;
  Timed tm(h, DGPU_K_ZINV, (2.0 * sizeof(R) * g.tp + 16.0 * g.sp) * (j1 - j0 + 1) * nf,
           2.0 * sizeof(R) * g.tp * (j1 - j0 + 1) * nf);
  if (h->zgen) {
    zgen_inv<R>(h, C, T, j0, j1, OP_ID, s0, s1, fb);
    return;
  }
  const int ng = z_ngroups<R>(h);
  dim3 grid((g.nq + ng - 1) / ng, j1 - j0 + 1, nf), block(ng * (g.nz / 8));
  auto kern = g.nz <= 512 ? k_zinv<R, true> : k_zinv<R, false>;
  if constexpr (sizeof(R) == 8)
    if (g.nz == 512 && (h->kn.fft_const & 1)) kern = k_zinv<R, true, 9>;
  const bool twp = g.nz == 512 && use_xl<R>(h, XL_ZINV) && xl_powers<R>(h, XL_ZINV);
  if (g.nz == 512 && use_xl<R>(h, XL_ZINV)) kern = twp ? k_zinv<R, true, FFT_XL512P> : k_zinv<R, true, FFT_XL512>;
  const size_t lds = z_lds<R>(h, ng, !twp);
  hipLaunchKernelGGL(kern, grid, block, lds, h->stream, reinterpret_cast<const cpx<double>*>(C), T, g, twz_of<R>(h),
                     j0, s0, s1, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

template <typename R>
void launch_zinv_op(dgpu_handle* h, const double2* C, cpx<R>* T, int j0, int j1, int op, double s0 = 1.0,
                    double s1 = 1.0, const FieldBatch* fb = nullptr) {
  if (j1 < j0) return;
  const Geo& g = h->g;
  const int nf = fb ? fb->nf : 1;
  Timed tm(h, DGPU_K_ZINV, (2.0 * sizeof(R) * g.tp + 16.0 * g.sp) * (j1 - j0 + 1) * nf,
           2.0 * sizeof(R) * g.tp * (j1 - j0 + 1) * nf);
  if (h->zgen) {
    zgen_inv<R>(h, C, T, j0, j1, op, s0, s1, fb);
    return;
  }
  const int ng = z_ngroups<R>(h);
  dim3 grid((g.nq + ng - 1) / ng, j1 - j0 + 1, nf), block(ng * (g.nz / 8));
  auto kern = g.nz <= 512 ? k_zinv_op<R, true> : k_zinv_op<R, false>;
  const bool twp = g.nz == 512 && use_xl<R>(h, XL_ZINV) && xl_powers<R>(h, XL_ZINV);
  if (g.nz == 512 && use_xl<R>(h, XL_ZINV)) kern = twp ? k_zinv_op<R, true, FFT_XL512P> : k_zinv_op<R, true, FFT_XL512>;
  const size_t lds = z_lds<R>(h, ng, !twp);
  hipLaunchKernelGGL(kern, grid, block, lds, h->stream, reinterpret_cast<const cpx<double>*>(C), T, g, twz_of<R>(h),
                     j0, op, h->d_kx.p, h->d_kz.p, h->d_dy.p, h->d_dyf.p, s0, s1, fb ? *fb : FieldBatch{});
  HIPCHECK(hipGetLastError());
}

// The CN solve when the rows depend on nu_t / kappa_t per column.
template <typename R>
void launch_thomas_les(dgpu_handle* h, int field, int rs, const R* r, R* u, const double* nt, int ca,
                       int cb, double transient, bool incr, const LesCnFields<R>* batch = nullptr) {
  const Geo& g = h->g;
  const int n = g.ny2;
  const TriRow* rows = h->trirows.p + ((size_t)field * 3 + rs) * n;
  const LesCnFields<R> mf = batch ? *batch : LesCnFields<R>{};
  const int nf = batch ? batch->nf : 1;
  // reads r, u, nu_t; writes the e, f scratch and reads it back; writes u.
  // Two columns per lane (one 2-wide access per row and array) when the
  // plane allows it.
  Timed tm(h, DGPU_K_THOMAS, (3.0 * sizeof(R) + 8.0 + 32.0) * g.pp * n * nf);
  if (h->kn.les_cols == 2 && g.pp % 2 == 0) {
    dim3 grid((unsigned)((g.pp / 2 + 255) / 256), (unsigned)nf), block(256);
    hipLaunchKernelGGL((k_thomas_les<R, 2>), grid, block, 0, h->stream, r, u, nt, h->sbl.p, h->sgl.p, rows, n,
                       g.pp, ca, cb, transient, incr ? 1 : 0, mf);
  } else {
    dim3 grid((unsigned)((g.pp + 255) / 256), (unsigned)nf), block(256);
    hipLaunchKernelGGL((k_thomas_les<R, 1>), grid, block, 0, h->stream, r, u, nt, h->sbl.p, h->sgl.p, rows, n,
                       g.pp, ca, cb, transient, incr ? 1 : 0, mf);
  }
  HIPCHECK(hipGetLastError());
}

// Planes per batch of a 2-D transform (see Knobs::batch_mb).
template <typename R>
int batch_planes(const dgpu_handle* h) {
  if (h->kn.batch_mb <= 0) return 1 << 30;
  const double per = 2.0 * sizeof(R) * (double)h->g.tp;
  int b = (int)((double)h->kn.batch_mb * 1048576.0 / per);
  return b < 8 ? 8 : b;
}

// spectral -> physical for one field, all planes (fft_xz_to_physical, fft.f90:319-382)
template <typename R>
void to_physical(dgpu_handle* h, const double2* C, R* u, cpx<R>* T) {
  const int n2 = h->g.ny2, B = batch_planes<R>(h);
  for (int j0 = 0; j0 < n2; j0 += B) {
    const int j1 = std::min(n2 - 1, j0 + B - 1);
    launch_zinv<R>(h, C, T, j0, j1);
    launch_xinv<R>(h, T, u, j0, j1);
  }
}

// spectral -> physical for several fields (C[k] -> u[k] through the x-spectra
// T[k]): every field's z pass in one launch, then every field's x pass.
template <typename R>
void to_physical_fields(dgpu_handle* h, int nf, const double2* const* C, R* const* u, cpx<R>* const* T) {
  if (batch_planes<R>(h) < h->g.ny2 || nf > MAXBF) {  // plane batches: field by field
    for (int k = 0; k < nf; ++k) to_physical<R>(h, C[k], u[k], T[k]);
    return;
  }
  FieldBatch zb{}, xb{};
  for (int k = 0; k < nf; ++k) {
    zb.in[k] = C[k];
    zb.out[k] = T[k];
    xb.in[k] = T[k];
    xb.out[k] = u[k];
  }
  zb.nf = xb.nf = nf;
  const int n2 = h->g.ny2;
  launch_zinv<R>(h, C[0], T[0], 0, n2 - 1, 1.0, 1.0, &zb);
  launch_xinv<R>(h, T[0], u[0], 0, n2 - 1, nullptr, &xb);
}

// x inverse of several fields' x-spectra (T[k] -> u[k], planes [j0, j1], each
// times i kx first with dkx), up to MAXBF fields a launch: one launch moves a
// batch at the rate of the whole-state pass instead of a field's at a time
template <typename R>
void xinv_fields(dgpu_handle* h, int nf, cpx<R>* const* T, R* const* u, int j0, int j1, const double* dkx = nullptr) {
  for (int k0 = 0; k0 < nf; k0 += MAXBF) {
    FieldBatch xb{};
    xb.nf = std::min(MAXBF, nf - k0);
    for (int k = 0; k < xb.nf; ++k) {
      xb.in[k] = T[k0 + k];
      xb.out[k] = u[k0 + k];
    }
    launch_xinv<R>(h, T[k0], u[k0], j0, j1, dkx, &xb);
  }
}

// physical -> spectral for one field on planes [j0, j1] (fft_xz_to_fourier,
// fft.f90:247-315); with accum the transform is added to C
template <typename R>
void to_fourier(dgpu_handle* h, const R* u, double2* C, cpx<R>* T, int j0, int j1, bool accum = false) {
  const int B = batch_planes<R>(h);
  for (int a = j0; a <= j1; a += B) {
    const int b = std::min(j1, a + B - 1);
    launch_xfwd<R>(h, u, T, a, b);
    launch_zfwd<R>(h, T, C, a, b, accum);
  }
}

// physical -> spectral for several fields on planes [j0, j1] (u[k] -> C[k]
// through the x-spectra T[k]): every field's x pass in one launch, then every
// field's z pass.
template <typename R>
void to_fourier_fields(dgpu_handle* h, int nf, R* const* u, double2* const* C, cpx<R>* const* T, int j0, int j1,
                       bool accum = false) {
  for (int k0 = 0; k0 < nf; k0 += MAXBF) {
    const int nb = std::min(MAXBF, nf - k0);
    if (batch_planes<R>(h) < h->g.ny2) {  // plane batches: field by field
      for (int k = k0; k < k0 + nb; ++k) to_fourier<R>(h, u[k], C[k], T[k], j0, j1, accum);
      continue;
    }
    FieldBatch xb{}, zb{};
    for (int k = 0; k < nb; ++k) {
      xb.in[k] = u[k0 + k];
      xb.out[k] = T[k0 + k];
      zb.in[k] = T[k0 + k];
      zb.out[k] = C[k0 + k];
    }
    xb.nf = zb.nf = nb;
    launch_xfwd<R>(h, u[k0], T[k0], j0, j1, &xb);
    launch_zfwd<R>(h, T[k0], C[k0], j0, j1, accum, 1.0, 1.0, &zb);
  }
}

// The CN solve of one field on planes [ca, cb]: physical columns (r, u with
// plane pitch pp), or with spec the dealiased spectra (r = cr, u = the
// spectral state, FP64, 2 sp real columns a plane).
// With r32 (spectral FP64 state only) r holds FP32 increments cr - u and
// the solve adds the state back in FP64 (k_thomas_seg's RPLUS).
template <typename R>
void launch_thomas(dgpu_handle* h, int field, int rs, const R* r, R* u, int ca, int cb, double transient,
                   bool incr, const SegRow* table = nullptr, bool spec = false,
                   const CnFields<R>* batch = nullptr, bool r32 = false) {
  const Geo& g = h->g;
  const int n = g.ny2;
  const SegRow* rows = (table ? table : h->rows.p) + ((size_t)field * 3 + rs) * n;
  const CnFields<R> mf = batch ? *batch : CnFields<R>{};
  const int nf = batch ? batch->nf : 1;
  const long long pitch = spec ? 2 * g.sp : g.pp;
  const long long ncol = pitch, nbc = spec ? 1 : 0;  // lanes past ncol (a partial last tile) touch no memory
  // columns per lane: Knobs::thomas_cols for segments of up to 32 rows (register budget), else 1.
  // The wider row runs (64 C values) pay for the lower occupancy: at 512x257x512 the FP64
  // solve takes 4.34 ms per step with two columns and 4.70 with one.
  const int C = (h->seg_L <= 32 && h->kn.thomas_cols == 2 && (spec || g.pp % 128 == 0)) ? 2 : 1;
  dim3 grid((unsigned)((pitch + 64 * C - 1) / (64 * C)), (unsigned)nf), block(64, h->seg_S);
  Timed tm(h, DGPU_K_THOMAS, (2.0 * sizeof(R) + (r32 ? 4.0 : sizeof(R))) * pitch * n * nf);
#define TLAUNCH(L, CC)                                                                                     \
  do {                                                                                                     \
    if (r32) {                                                                                             \
      if constexpr (sizeof(R) == 8)                                                                        \
        hipLaunchKernelGGL((k_thomas_seg<L, CC, double, false, false, float, true>), grid, block, 0,       \
                           h->stream, reinterpret_cast<const float*>(r), u, rows, n, h->seg_S, pitch, ca,  \
                           cb, transient, nullptr, nullptr, nullptr, ncol, nbc, mf);                      \
    } else if (incr)                                                                                       \
      hipLaunchKernelGGL((k_thomas_seg<L, CC, R, true>), grid, block, 0, h->stream, r, u, rows, n,         \
                         h->seg_S, pitch, ca, cb, transient, nullptr, nullptr, nullptr, ncol, nbc, mf);   \
    else                                                                                                   \
      hipLaunchKernelGGL((k_thomas_seg<L, CC, R, false>), grid, block, 0, h->stream, r, u, rows, n,        \
                         h->seg_S, pitch, ca, cb, transient, nullptr, nullptr, nullptr, ncol, nbc, mf);   \
  } while (0)
  switch (h->seg_L * 4 + C) {
    case 8 * 4 + 1: TLAUNCH(8, 1); break;
    case 16 * 4 + 1: TLAUNCH(16, 1); break;
    case 17 * 4 + 1: TLAUNCH(17, 1); break;
    case 20 * 4 + 1: TLAUNCH(20, 1); break;
    case 24 * 4 + 1: TLAUNCH(24, 1); break;
    case 28 * 4 + 1: TLAUNCH(28, 1); break;
    case 32 * 4 + 1: TLAUNCH(32, 1); break;
    case 40 * 4 + 1: TLAUNCH(40, 1); break;
    case 48 * 4 + 1: TLAUNCH(48, 1); break;
    case 56 * 4 + 1: TLAUNCH(56, 1); break;
    case 8 * 4 + 2: TLAUNCH(8, 2); break;
    case 16 * 4 + 2: TLAUNCH(16, 2); break;
    case 17 * 4 + 2: TLAUNCH(17, 2); break;
    case 20 * 4 + 2: TLAUNCH(20, 2); break;
    case 24 * 4 + 2: TLAUNCH(24, 2); break;
    case 28 * 4 + 2: TLAUNCH(28, 2); break;
    case 32 * 4 + 2: TLAUNCH(32, 2); break;
    default:
      throw ApiError{DGPU_EUNSUP, "unsupported Thomas segment length"};
  }
#undef TLAUNCH
  HIPCHECK(hipGetLastError());
}

// The adjoint's transposed CN solve vbar = A^-T xbar (factor table rowsT) with
// its epilogue: rbar = vbar on the assembled rows, ubar = E^T vbar.  Physical
// columns (xb, rbar, ubar with plane pitch pp, all rows), or with spec the
// spectral adjoints: xb == ubar is the state adjoint, solved on its planes
// [ca, cb] and updated in place, rbar the adjoint of cr (2 sp real columns a
// plane).
void launch_thomas_adj(dgpu_handle* h, int field, int rs, const double* xb, double* rbar, double* ubar,
                       bool spec = false, int ca = 0, int cb = -1) {
  const Geo& g = h->g;
  const int n = g.ny2;
  const size_t off = ((size_t)field * 3 + rs) * n;
  const SegRow* rows = h->rowsT.p + off;
  const SegRow* prow = h->rows.p + off;
  const long long pitch = spec ? 2 * g.sp : g.pp;
  const long long ncol = pitch;
  if (!spec) {
    ca = 0;
    cb = n - 1;
  }
  const int C = (h->seg_L <= 32 && h->kn.thomas_cols == 2 && (spec || g.pp % 128 == 0)) ? 2 : 1;
  dim3 grid((unsigned)((pitch + 64 * C - 1) / (64 * C))), block(64, h->seg_S);
  Timed tm(h, DGPU_K_THOMAS, 24.0 * pitch * n);
#define TLAUNCH(L, CC)                                                                                     \
  hipLaunchKernelGGL((k_thomas_seg<L, CC, double, false, true>), grid, block, 0, h->stream, xb,             \
                     (double*)nullptr, rows, n, h->seg_S, pitch, ca, cb, 0.0, prow, rbar, ubar, ncol, 0LL)
  switch (h->seg_L * 4 + C) {
    case 8 * 4 + 1: TLAUNCH(8, 1); break;
    case 16 * 4 + 1: TLAUNCH(16, 1); break;
    case 17 * 4 + 1: TLAUNCH(17, 1); break;
    case 20 * 4 + 1: TLAUNCH(20, 1); break;
    case 24 * 4 + 1: TLAUNCH(24, 1); break;
    case 28 * 4 + 1: TLAUNCH(28, 1); break;
    case 32 * 4 + 1: TLAUNCH(32, 1); break;
    case 40 * 4 + 1: TLAUNCH(40, 1); break;
    case 48 * 4 + 1: TLAUNCH(48, 1); break;
    case 56 * 4 + 1: TLAUNCH(56, 1); break;
    case 8 * 4 + 2: TLAUNCH(8, 2); break;
    case 16 * 4 + 2: TLAUNCH(16, 2); break;
    case 17 * 4 + 2: TLAUNCH(17, 2); break;
    case 20 * 4 + 2: TLAUNCH(20, 2); break;
    case 24 * 4 + 2: TLAUNCH(24, 2); break;
    case 28 * 4 + 2: TLAUNCH(28, 2); break;
    case 32 * 4 + 2: TLAUNCH(32, 2); break;
    default:
      throw ApiError{DGPU_EUNSUP, "unsupported Thomas segment length"};
  }
#undef TLAUNCH
  HIPCHECK(hipGetLastError());
}

void launch_herm0(dgpu_handle* h, double2* c, int ca, int cb) {
  if (cb < ca) return;
  hipLaunchKernelGGL(k_herm0, dim3((unsigned)((h->g.nkz + 64) / 64), (unsigned)(cb - ca + 1)), dim3(64), 0,
                     h->stream, c, h->g, ca, cb);
  HIPCHECK(hipGetLastError());
}

void copy_spec(dgpu_handle* h, double2* dst, const double2* src) {
  const long long n = (long long)h->g.ny2 * h->g.sp;
  Timed tm(h, DGPU_K_COPY, 32.0 * n);
  hipLaunchKernelGGL(k_copy2, dim3(2048), dim3(256), 0, h->stream, dst, src, n);
  HIPCHECK(hipGetLastError());
}

void launch_remdiv(dgpu_handle* h, bool update_cp, double temp4) {
  const Geo& g = h->g;
  RemDivArgs a;
  a.cu1 = h->cu[0].p;
  a.cu2 = h->cu[1].p;
  a.cu3 = h->cu[2].p;
  a.cp = h->cp.p;
  a.ib = h->rd_ib.p;
  a.kx = h->d_kx.p;
  a.kz = h->d_kz.p;
  a.inv_h = 1.0 / temp4;
  a.update_cp = update_cp ? 1 : 0;
  a.js = h->jstart;
  a.je = h->jend;
  // k_remdiv_seg addresses a field with 32-bit byte offsets
  if ((double)g.ny2 * (double)g.sp * 16.0 >= 4294967295.0)
    throw ApiError{DGPU_EUNSUP, "spectral field too large for the projection"};
  Timed tm(h, DGPU_K_REMDIV, (update_cp ? 128.0 : 96.0) * g.sp * g.ny);
  // rows 2..Ny in segments; the block holds whole wavefronts (64 / MW
  // segments each), the segments past Ny being empty
  const int nr = g.ny - 1;
  auto launch = [&](auto kern, int L, int MW) {
    const int spw = 64 / MW;
    const int S = ((nr + L - 1) / L + spw - 1) / spw * spw;
    hipLaunchKernelGGL(kern, dim3((unsigned)((g.sp + MW - 1) / MW)), dim3(MW * S), 0, h->stream, a, g,
                       h->d_prow.p, S);
  };
  if (nr <= 4 * 64)
    launch(k_remdiv_seg<4, 16, 1024>, 4, 16);
  else if (nr <= 8 * 64)
    launch(k_remdiv_seg<8, 8, 512>, 8, 8);
  else if (nr <= 8 * 128)
    launch(k_remdiv_seg<8, 4, 512>, 8, 4);
  else
    throw ApiError{DGPU_EUNSUP, "ny too large for the projection"};
  HIPCHECK(hipGetLastError());
}

// The projection's row coefficients and the pivots of every mode (once per
// handle: they depend only on the grid and the wavenumbers).
void setup_projection(dgpu_handle* h) {
  const Geo& g = h->g;
  const int ny = g.ny;
  std::vector<PoisRow> rows((size_t)ny + 2, PoisRow{0.0, 0.0, 0.0, 0.0});
  for (int j = 1; j <= ny; ++j) {
    PoisRow& r = rows[j];
    if (j >= 2 && j < ny) {
      r.ml = 1.0 / (h->dy[j] * h->dyf[j]);
      r.mu = 1.0 / (h->dy[j + 1] * h->dyf[j]);
    } else if (j == ny) {
      r.ml = 1.0;
      r.mu = 0.0;
    }
    r.idyf = 1.0 / h->dyf[j];
    r.idy = (j >= 2) ? 1.0 / h->dy[j] : 0.0;
  }
  upload(h->d_prow, rows);
  h->rd_ib.alloc((size_t)g.ny2 * piv_plane(g));  // kz >= 0: the -kz modes share the pivots
  hipLaunchKernelGGL(k_remdiv_pivots, dim3((unsigned)((piv_plane(g) + 255) / 256)), dim3(256), 0, h->stream,
                     h->rd_ib.p, h->d_prow.p, h->d_kx.p, h->d_kz.p, g);
  HIPCHECK(hipGetLastError());
}

void launch_bc_post(dgpu_handle* h, double time) {
  const dgpu_params& p = h->p;
  BcPostArgs a;
  a.cu1 = h->cu[0].p;
  a.cu2 = h->cu[1].p;
  a.cu3 = h->cu[2].p;
  a.dy = h->d_dy.p;
  a.dyf = h->d_dyf.p;
  a.u_lo = p.u_bc_ymin;
  a.v_lo = p.v_bc_ymin;
  a.w_lo = p.w_bc_ymin;
  a.u_hi = p.u_bc_ymax;
  a.v_hi = p.v_bc_ymax;
  a.w_hi = p.w_bc_ymax;
  a.u_lo_c1 = p.u_bc_ymin_c1;
  a.v_lo_c1 = p.v_bc_ymin_c1;
  a.w_lo_c1 = p.w_bc_ymin_c1;
  a.u_hi_c1 = p.u_bc_ymax_c1;
  a.v_hi_c1 = p.v_bc_ymax_c1;
  a.w_hi_c1 = p.w_bc_ymax_c1;
  if (p.w_bc_ymax == 1 && p.f_type == 4)
    h->w_transient =
        p.w_bc_ymax_c1 + p.amp_omega0 * std::sin(p.omega0 * (p.ro_inv / p.delta * time - p.force_start));
  a.w_hi_c1_transient = h->w_transient;
  dim3 grid((unsigned)((h->g.sp + 255) / 256)), block(256);
  Timed tm(h, DGPU_K_BCPOST, 16.0 * 14 * h->g.sp);
  hipLaunchKernelGGL(k_bc_post, grid, block, 0, h->stream, a, h->g);
  HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------------ LES --
// les_chan + les_chan_th (les.f90:20-1455) for the substep: wall and ghost
// values on the spectral state, the nine velocity gradients (and three per
// scalar) to physical space, the model kernels, and the forward transforms
// and divergence that give the subgrid forcing of cf1..3 and cfth.  Leaves
// u1..3 and th in physical space, as the reference's les_chan does.
template <typename R>
void les_stage(dgpu_handle* h, double time) {
  const dgpu_params& p = h->p;
  const Geo& g = h->g;
  const int nth = h->nth, Ny = g.ny, n2 = g.ny2;
  launch_bc_post(h, time);  // apply_BC_vel_mpi_post (les.f90:80)
  {
    LesBcArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cu1 = h->cu[0].p;
    a.cu3 = h->cu[2].p;
    a.dy = h->d_dy.p;
    a.nth = nth;
    a.u_lo = p.u_bc_ymin;
    a.u_hi = p.u_bc_ymax;
    a.w_lo = p.w_bc_ymin;
    a.w_hi = p.w_bc_ymax;
    for (int n = 0; n < nth; ++n) {
      a.cth[n] = h->cth[n].p;
      a.th_lo[n] = p.th_bc_ymin[n];
      a.th_hi[n] = p.th_bc_ymax[n];
      a.th_lo_c1[n] = p.th_bc_ymin_c1[n];
      a.th_hi_c1[n] = p.th_bc_ymax_c1[n];
    }
    dim3 grid((unsigned)((g.sp + 255) / 256)), block(256);
    Timed tm(h, DGPU_K_BCPOST, 16.0 * 12 * g.sp * (1 + nth));
    hipLaunchKernelGGL(k_les_bc, grid, block, 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  // gradients to physical space
  cpx<R>* T = xspec<R>(h->tpool[0]);
  const double2* cu1 = h->cu[0].p;
  const double2* cu2 = h->cu[1].p;
  const double2* cu3 = h->cu[2].p;
  const struct {
    const double2* c;
    int op;
  } gsrc[9] = {{cu1, OP_DX}, {cu1, OP_DY_GY}, {cu1, OP_DZ},  {cu2, OP_DX}, {cu2, OP_DY_GYF},
               {cu2, OP_DZ}, {cu3, OP_DX},    {cu3, OP_DY_GY}, {cu3, OP_DZ}};
  // the state itself (les.f90:586-588, 1207-1209, 1229-1231); its x-spectra
  // stay in tpool[f] (all planes, without plane batches) for the x derivatives
  {
    const double2* C[3 + DGPU_MAX_TH];
    R* U[3 + DGPU_MAX_TH];
    cpx<R>* T[3 + DGPU_MAX_TH];
    for (int f = 0; f < 3 + nth; ++f) {
      C[f] = f < 3 ? h->cu[f].p : h->cth[f - 3].p;
      U[f] = f < 3 ? phys<R>(h->u[f]) : phys<R>(h->th[f - 3]);
      T[f] = xspec<R>(h->tpool[f]);
    }
    to_physical_fields<R>(h, 3 + nth, C, U, T);
  }
  // x derivatives: i kx times the state's x-spectra, then the x inverse; z
  // derivatives: i kz inside the z inverse; the y differences (G12, G22,
  // G32, the scalars' T2) are formed from the physical state in k_les_model
  const bool share = h->kn.batch_mb <= 0;
  if (share) {
    // the x derivatives of every field in one launch from the state's
    // x-spectra; then the z derivatives' x-spectra reuse those buffers and
    // go through one more batched x inverse
    const int nd = 3 + nth;
    cpx<R>* TX[3 + DGPU_MAX_TH];
    R* DX[3 + DGPU_MAX_TH];
    R* DZ[3 + DGPU_MAX_TH];
    for (int f = 0; f < nd; ++f) {
      TX[f] = xspec<R>(h->tpool[f]);
      DX[f] = f < 3 ? phys<R>(h->grad[3 * f]) : phys<R>(h->dth[0][f - 3]);
      DZ[f] = f < 3 ? phys<R>(h->grad[3 * f + 2]) : phys<R>(h->dth[2][f - 3]);
    }
    xinv_fields<R>(h, nd, TX, DX, 0, n2 - 1, h->d_kx.p);
    for (int f0 = 0; f0 < nd; f0 += MAXBF) {
      FieldBatch zb{};
      zb.nf = std::min(MAXBF, nd - f0);
      for (int k = 0; k < zb.nf; ++k) {
        const int f = f0 + k;
        zb.in[k] = f < 3 ? h->cu[f].p : h->cth[f - 3].p;
        zb.out[k] = TX[f];
      }
      launch_zinv_op<R>(h, static_cast<const double2*>(zb.in[0]), TX[f0], 0, n2 - 1, OP_DZ, 1.0, 1.0, &zb);
    }
    xinv_fields<R>(h, nd, TX, DZ, 0, n2 - 1);
  }
  T = xspec<R>(h->tpool[3 + nth]);  // scratch past the state's x-spectra
  for (int m = 0; !share && m < 9; ++m) {
    if (gsrc[m].op == OP_DY_GY || gsrc[m].op == OP_DY_GYF) continue;
    launch_zinv_op<R>(h, gsrc[m].c, T, 0, n2 - 1, gsrc[m].op);
    launch_xinv<R>(h, T, phys<R>(h->grad[m]), 0, n2 - 1);
  }
  for (int n = 0; !share && n < nth; ++n) {
    launch_zinv_op<R>(h, h->cth[n].p, T, 0, n2 - 1, OP_DX);
    launch_xinv<R>(h, T, phys<R>(h->dth[0][n]), 0, n2 - 1);
    launch_zinv_op<R>(h, h->cth[n].p, T, 0, n2 - 1, OP_DZ);
    launch_xinv<R>(h, T, phys<R>(h->dth[2][n]), 0, n2 - 1);
  }
  // velocity and scalar models in one pass over the points
  {
    LesVelArgs<R> a;
    for (int m = 0; m < 9; ++m) a.G[m] = phys<R>(h->grad[m]);
    a.u1 = phys<R>(h->u[0]);
    a.u3 = phys<R>(h->u[2]);
    a.tau11 = phys<R>(h->tau[0]);
    a.tau33 = phys<R>(h->tau[1]);
    a.tau13 = phys<R>(h->tau[2]);
    a.tau12 = phys<R>(h->tau[3]);
    a.tau23 = phys<R>(h->tau[4]);
    a.cr1 = phys<R>(h->tau[5]);
    a.cr3 = phys<R>(h->tau[6]);
    a.nu_t = h->nu_t.p;
    a.dy = h->d_dy.p;
    a.dyf = h->d_dyf.p;
    a.dx1 = p.lx / p.nx;
    a.dz1 = p.lz / p.nz;
    a.type = p.les_model_type;
    a.js = h->jstart;
    a.je = h->jend;
    LesThAll<R> th;
    std::memset(&th, 0, sizeof(th));
    th.nth = nth;
    for (int n = 0; n < nth; ++n) {
      LesThArgs<R>& t = th.s[n];
      for (int m = 0; m < 9; ++m) t.G[m] = phys<R>(h->grad[m]);
      t.T1 = phys<R>(h->dth[0][n]);
      t.T2 = (p.les_model_type == 5) ? phys<R>(h->dth[1][n]) : phys<R>(h->dth[0][n]);
      t.T3 = phys<R>(h->dth[2][n]);
      t.nu_t = h->nu_t.p;
      t.f1 = phys<R>(h->thf[0][n]);
      t.f3 = phys<R>(h->thf[1][n]);
      t.kappa_t = h->kappa_t[n].p;
      t.dy = h->d_dy.p;
      t.dyf = h->d_dyf.p;
      t.dx1 = p.lx / p.nx;
      t.dz1 = p.lz / p.nz;
      t.type = p.les_model_type;
      t.jst = h->jstart_th[n];
      t.jet = h->jend_th[n];
    }
    LesDyArgs<R> dy;
    std::memset(&dy, 0, sizeof(dy));
    dy.u2 = phys<R>(h->u[1]);
    for (int n = 0; n < nth; ++n) {
      dy.th[n] = phys<R>(h->th[n]);
      dy.t2[n] = phys<R>(h->dth[1][n]);
    }
    dy.g12 = phys<R>(h->grad[1]);
    dy.g22 = phys<R>(h->grad[4]);
    dy.g32 = phys<R>(h->grad[7]);
    dy.store = h->les_cap ? 1 : 0;  // the adjoint's gathers read them
    const long long npt = g.pp * (Ny + 1);
    dim3 grid((unsigned)((npt + 255) / 256)), block(256);
    // 6 gradients, u1, u2, u3 in; 5 stresses, 2 fluxes, nu_t out; per scalar
    // th and 2 gradients in, 2 fluxes and kappa_t out
    Timed tm(h, DGPU_K_LES,
             ((6.0 + 3.0 + 7.0) * sizeof(R) + 8.0 + nth * ((3.0 + 2.0) * sizeof(R) + 8.0)) * npt);
    hipLaunchKernelGGL(k_les_model<R>, grid, block, 0, h->stream, a, th, dy, g);
    HIPCHECK(hipGetLastError());
  }
  {  // the stresses and fluxes to Fourier space, batched over fields (the
     // x-spectra buffers are idle until the products pass)
    R* U[7 + 2 * DGPU_MAX_TH];
    double2* C[7 + 2 * DGPU_MAX_TH];
    cpx<R>* T[7 + 2 * DGPU_MAX_TH];
    int nf = 0;
    for (int m = 0; m < 7; ++m, ++nf) {
      U[nf] = phys<R>(h->tau[m]);
      C[nf] = h->ctau[m].p;
    }
    for (int n = 0; n < nth; ++n)
      for (int w = 0; w < 2; ++w, ++nf) {
        U[nf] = phys<R>(h->thf[w][n]);
        C[nf] = h->cthf[w][n].p;
      }
    for (int k = 0; k < nf; ++k) T[k] = xspec<R>(h->tpool[k % (int)h->tpool.size()]);
    to_fourier_fields<R>(h, nf, U, C, T, 1, Ny);
  }
  // divergence into the subgrid forcing
  {
    LesDivArgs a;
    std::memset(&a, 0, sizeof(a));
    a.t11 = h->ctau[0].p;
    a.t33 = h->ctau[1].p;
    a.t13 = h->ctau[2].p;
    a.t12 = h->ctau[3].p;
    a.t23 = h->ctau[4].p;
    a.c1 = h->ctau[5].p;
    a.c3 = h->ctau[6].p;
    a.f1 = h->lesf[0].p;
    a.f2 = h->lesf[1].p;
    a.f3 = h->lesf[2].p;
    for (int n = 0; n < nth; ++n) {
      a.th_f1[n] = h->cthf[0][n].p;
      a.th_f3[n] = h->cthf[1][n].p;
      a.fth[n] = h->lesfth[n].p;
      a.jst[n] = h->jstart_th[n];
      a.jet[n] = h->jend_th[n];
    }
    a.kx = h->d_kx.p;
    a.kz = h->d_kz.p;
    a.dyf = h->d_dyf.p;
    a.js = h->jstart;
    a.j2 = Ny - 1;
    a.nth = nth;
    a.th_sign = (p.les_model_type == 5) ? -1 : 1;
    const long long nm = g.sp * n2;
    dim3 grid((unsigned)((nm + 255) / 256)), block(256);
    Timed tm(h, DGPU_K_LES, 16.0 * (7 + 3 + 3 * nth) * nm);
    hipLaunchKernelGGL(k_les_div, grid, block, 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
}

// ------------------------------------------------------ poisson_p_chan --
// cp on planes 1..Ny from cu1..3 and cth (channel.f90:1026-1234).  The nine
// derivative fields go to physical space one or two at a time (scratch r[1],
// r[2]), their products accumulate in r[0], one forward transform gives the
// spectral right-hand side, and k_poisson_solve adds the buoyancy term and
// solves every mode.
void poisson_p(dgpu_handle* h, const double* ri) {
  const Geo& g = h->g;
  const int j0 = 2, j1 = g.ny;
  if (!h->pois.p) h->pois.alloc((size_t)g.ny2 * g.sp);
  double* acc = h->r[0].p;
  double* pa = h->r[1].p;
  double* pb = h->r[2].p;
  cpx<double>* T = xspec<double>(h->tpool[0]);
  auto phys = [&](int f, int op, double* out) {
    launch_zinv_op<double>(h, h->cu[f].p, T, j0, j1, op);
    launch_xinv<double>(h, T, out, j0, j1);
  };
  const long long npts = (long long)(j1 - j0 + 1) * g.nz * g.nx;
  auto accum = [&](int mode, const double* a, const double* b) {
    const unsigned nb = (unsigned)std::min<long long>(4096, (npts + 255) / 256);
    Timed tm(h, DGPU_K_INIT, (mode == 0 ? 16.0 : mode == 1 ? 24.0 : 32.0) * npts);
    hipLaunchKernelGGL(k_poisson_accum, dim3(nb), dim3(256), 0, h->stream, acc, a, b, mode, j0, j1, g.nx, g.nz,
                       (long long)g.pp);
    HIPCHECK(hipGetLastError());
  };
  phys(0, OP_DX, pa);  // (du/dx)^2, (dv/dy)^2, (dw/dz)^2 (channel.f90:1041-1072)
  accum(0, pa, nullptr);
  phys(1, OP_DY_GYF, pa);
  accum(1, pa, nullptr);
  phys(2, OP_DZ, pa);
  accum(1, pa, nullptr);
  phys(0, OP_DY_C2, pa);  // 2 (du/dy)(dv/dx) (channel.f90:1075-1106)
  phys(1, OP_DX_AVG, pb);
  accum(2, pa, pb);
  phys(2, OP_DY_C2, pa);  // 2 (dw/dy)(dv/dz) (channel.f90:1108-1140)
  phys(1, OP_DZ_AVG, pb);
  accum(2, pa, pb);
  phys(0, OP_DZ, pa);  // 2 (du/dz)(dw/dx) (channel.f90:1142-1172)
  phys(2, OP_DX, pb);
  accum(2, pa, pb);
  to_fourier<double>(h, acc, h->pois.p, T, j0, j1);
  PoissonArgs a;
  std::memset(&a, 0, sizeof(a));
  a.rhs = h->pois.p;
  a.nth = h->nth;
  for (int n = 0; n < h->nth; ++n) {
    a.cth[n] = h->cth[n].p;
    a.ri[n] = ri[n];
  }
  a.kx = h->d_kx.p;
  a.kz = h->d_kz.p;
  a.dy = h->d_dy.p;
  a.dyf = h->d_dyf.p;
  a.gyf = h->d_gyf.p;
  a.sb = h->sb.p;
  a.sg = h->sg.p;
  a.cp = h->cp.p;
  Timed tm(h, DGPU_K_INIT, (48.0 + 16.0 * h->nth + 16.0) * g.sp * g.ny);
  hipLaunchKernelGGL(k_poisson_solve, dim3((unsigned)((g.sp + 255) / 256)), dim3(256), 0, h->stream, a, g);
  HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------- courant --
// The smallest per-point limit of courant over planes [j0, j1] of the given
// velocity buffers (physical, or spectral read as in pre_first_step).
double cfl_min(dgpu_handle* h, const void* u0, const void* u1, const void* u2, bool spectral) {
  const Geo& g = h->g;
  const dgpu_params& p = h->p;
  const bool front = h->cour.front != 0;
  const int j0 = front ? h->jstart : 1, j1 = front ? h->jend : g.ny;
  if (h->cfl_nset == 0 || j1 < j0) return HUGE_VAL;
  CflArgs a;
  std::memset(&a, 0, sizeof(a));
  a.u[0] = u0;
  a.u[1] = u1;
  a.u[2] = u2;
  a.spectral = spectral ? 1 : 0;
  a.cdy = h->cfl_cdy.p;
  a.cdx = h->cour.cfl * (p.lx / p.nx);  // dx = Lx/Nx (domain.f90:218)
  a.cdz = h->cour.cfl * (p.lz / p.nz);  // dz = Lz/Nz (domain.f90:260)
  a.offx = h->cfl_offx.p;
  a.offz = h->cfl_offz.p;
  a.nset = h->cfl_nset;
  a.j0 = j0;
  a.j1 = j1;
  a.partial = h->cfl_part.p;
  const long long n = (long long)(j1 - j0 + 1) * g.nz * g.nx;
  const int nb = (int)std::min<long long>((long long)h->cfl_part.n, (n + 255) / 256);
  {
    Timed tm(h, DGPU_K_INIT, 24.0 * n);
    hipLaunchKernelGGL(k_cfl_min, dim3((unsigned)nb), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  std::vector<double> part(nb);
  HIPCHECK(hipMemcpyAsync(part.data(), h->cfl_part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  double m = HUGE_VAL;
  for (double v : part) m = std::fmin(m, v);
  return m;
}

void build_adj_tables(dgpu_handle* h);

// delta_t, h_bar and the CN tables for a new time step.
void set_delta_t_impl(dgpu_handle* h, double dt) {
  if (!(dt > 0.0) || !std::isfinite(dt)) throw ApiError{DGPU_EINVAL, "delta_t must be positive and finite"};
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->p.delta_t = dt;
  h->h_bar[0] = dt * (8.0 / 15.0);  // advance.f90:154-156
  h->h_bar[1] = dt * (2.0 / 15.0);
  h->h_bar[2] = dt * (5.0 / 15.0);
  build_cn_tables(h);
  if (h->adj_ready) build_adj_tables(h);
}

// The scalar part of courant (advance.f90:62-158) given the smallest per-point
// limit `pts`: max_dt, the inertial and buoyancy periods, the caps, and the
// scheduling towards the next output.  pi is the single-precision value of
// fft.f90:152; the literals 0.1 and 1. - 0.025 are single precision there.
double courant_dt(const dgpu_handle* h, double pts, double time) {
  const dgpu_params& p = h->p;
  const dgpu_courant_params& c = h->cour;
  double dt = c.max_dt;
  if (p.ro_inv != 0.0) dt = std::min(dt, 2.0 * h->pi / std::fabs(p.ro_inv / p.delta) / 20.0);
  for (int n = 0; n < h->nth; ++n) {
    const double nmax = std::sqrt(std::fabs(p.th_bc_ymin_c1[n]));
    dt = std::min(dt, (double)0.1f * 2.0 * h->pi / nmax);
  }
  dt = std::min(dt, pts);
  double delta_t;
  if (dt <= 0.0)
    delta_t = 0.00001;
  else if (dt >= 1000.0)
    delta_t = 1000.0;
  else
    delta_t = dt;
  const double next = std::min(std::min(c.time_limit, c.save_stats_time - time),
                               std::min(c.save_flow_time - time, c.time_limit - time));
  const double r_max = (double)(1.0f - 0.025f);
  // the event scheduling below replaces the capped value whatever it was
  if (next < dt)
    delta_t = next + 1.0e-14;
  else if (next < (1.0 + r_max) * dt)
    delta_t = 1.0 / (1.0 + r_max) * next;
  else
    delta_t = dt;
  return delta_t;
}

void set_courant_impl(dgpu_handle* h, const dgpu_courant_params* c) {
  if (!c) {
    h->var_dt = false;
    return;
  }
  if (!(c->cfl > 0.0) || !(c->max_dt > 0.0) || c->update_dt < 1)
    throw ApiError{DGPU_EINVAL, "courant needs cfl > 0, max_dt > 0 and update_dt >= 1"};
  const dgpu_params& p = h->p;
  const int n2 = h->g.ny2;
  h->cour = *c;
  std::vector<double> cdy(n2, 0.0);
  for (int j = 1; j < n2; ++j) cdy[j] = c->cfl * h->dy[j];
  const int nset = c->front ? h->nth : 1;
  std::vector<double> offx((size_t)std::max(nset, 1) * n2, 0.0), offz(offx.size(), 0.0);
  if (c->front) {
    // thermal wind of the Front flavour (advance.f90:110-122)
    const double r = p.ro_inv / p.delta;
    for (int n = 0; n < nset; ++n)
      for (int j = 0; j < n2; ++j) {
        offx[(size_t)n * n2 + j] = p.dthdz[n] * h->gyf[j] / r;
        offz[(size_t)n * n2 + j] = (1.0 / r) * p.dthdx[n] * (h->gyf[j] - 0.5 * c->ly);
      }
  }
  upload(h->cfl_cdy, cdy);
  upload(h->cfl_offx, offx);
  upload(h->cfl_offz, offz);
  if (!h->cfl_part.p) h->cfl_part.alloc(1024);
  h->cfl_nset = nset;
  h->var_dt = true;
}

// ------------------------------------------------- state finiteness --
// Arrays of doubles scanned for NaN and Inf, one per blockIdx.y.
struct ScanList {
  const void* a[3 * (3 + DGPU_MAX_TH) + 1];
  long long n[3 * (3 + DGPU_MAX_TH) + 1];  // values
  int f32[3 * (3 + DGPU_MAX_TH) + 1];      // 1: FP32 values (the mixed form's RK carry)
};
static_assert(MAXTH == DGPU_MAX_TH, "the kernels' scalar cap is the ABI's");
constexpr int kScanBlocks = 128;  // blocks per array
constexpr int kScanFields = 3 * (3 + DGPU_MAX_TH) + 1;  // arrays per scan, at most

// Scratch of dgpu_check_state_noalloc: the scan's per-block flags in static
// device memory, so that the call allocates nothing.  Each GPU the library
// runs on holds one copy (29 KiB) until the process ends; every call on
// that GPU shares it, so calls that overlap on one GPU race.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (scratch_address); the name itself is a host stand-in.
constexpr int kScratchInts = kScanBlocks * kScanFields;
static __device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once
// per device (at dgpu_create).
void* scratch_address() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw ApiError{DGPU_EUNSUP, "device index beyond the scratch table"};
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw ApiError{DGPU_EHIP, "hipGetSymbolAddress returned no address for the scratch array"};
    scratch_ptr[dev] = ptr;
  }
  return scratch_ptr[dev];
}

// Each block writes one flag: 1 when its grid-stride share of the array
// holds a value that is not finite.  Every flag is written, so none needs
// clearing first.
__global__ void __launch_bounds__(256) k_nonfinite(ScanList L, int* __restrict__ flags) {
  const long long n = L.n[blockIdx.y];
  int bad = 0;
  if (L.f32[blockIdx.y]) {
    const float* a = static_cast<const float*>(L.a[blockIdx.y]);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
      bad |= !isfinite(a[i]);
  } else {
    const double* a = static_cast<const double*>(L.a[blockIdx.y]);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
      bad |= !isfinite(a[i]);
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) flags[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = bad;
}

// The finiteness check of dgpu_check_state: the spectral state, the RK
// carry and, with LES, the eddy viscosity and diffusivities.  flags holds
// kScanBlocks ints per scanned array on the device (the handle's buffer,
// allocated on first use, or the static scratch with noalloc), host the
// same on the host (a stack array with noalloc).
int check_state_impl(dgpu_handle* h, bool noalloc) {
  const Geo& g = h->g;
  ScanList L{};
  const char* names[kScanFields];
  int nf = 0;
  const long long ns = 2LL * g.sp * g.ny2;  // doubles of a spectral field
  auto add = [&](const void* p, long long n, const char* name, int f32 = 0) {
    if (!p) return;
    if (nf >= kScanFields) throw ApiError{DGPU_EINVAL, "more arrays than the scan holds"};
    L.a[nf] = p;
    L.n[nf] = n;
    L.f32[nf] = f32;
    names[nf++] = name;
  };
  const int c32 = h->precision == DGPU_MIXED ? 1 : 0;  // the carry's storage (carry32)
  static const char* thn[] = {"cth", "cfth", "kappa_t"};
  add(h->cu[0].p, ns, "cu1");
  add(h->cu[1].p, ns, "cu2");
  add(h->cu[2].p, ns, "cu3");
  add(h->cp.p, ns, "cp");
  add(h->cf[0].p, ns, "cf1", c32);
  add(h->cf[1].p, ns, "cf2", c32);
  add(h->cf[2].p, ns, "cf3", c32);
  for (int n = 0; n < h->nth; ++n) {
    add(h->cth[n].p, ns, thn[0]);
    add(h->cfth[n].p, ns, thn[1], c32);
  }
  if (h->p.use_les) {  // physical FP64 arrays in both forms
    const long long np = (long long)g.pp * g.ny2;
    add(h->nu_t.p, np, "nu_t");
    for (int n = 0; n < h->nth; ++n) add(h->kappa_t[n].p, np, thn[2]);
  }
  int* flags;
  int fl_stack[kScratchInts];
  std::vector<int> fl_vec;
  int* fl = fl_stack;
  if (noalloc) {
    // the static scratch of this handle's GPU, looked up at dgpu_create
    if (!h->scratch) throw ApiError{DGPU_EHIP, "no scratch address for this handle's device"};
    if ((size_t)kScanBlocks * nf > (size_t)kScratchInts)
      throw ApiError{DGPU_EINVAL, "scan larger than the scratch array (kScanBlocks x arrays)"};
    flags = static_cast<int*>(h->scratch);
  } else {
    if (!h->d_flags.p) h->d_flags.alloc((size_t)kScanBlocks * kScanFields);
    flags = h->d_flags.p;
    fl_vec.resize((size_t)kScanBlocks * nf);
    fl = fl_vec.data();
  }
  hipLaunchKernelGGL(k_nonfinite, dim3(kScanBlocks, nf), dim3(256), 0, h->stream, L, flags);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(fl, flags, (size_t)kScanBlocks * nf * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::string bad;
  for (int f = 0; f < nf; ++f)
    for (int b = 0; b < kScanBlocks; ++b)
      if (fl[(size_t)f * kScanBlocks + b]) {
        bad += bad.empty() ? names[f] : std::string(", ") + names[f];
        break;
      }
  if (!bad.empty()) throw ApiError{DGPU_ESTATE, "non-finite values in " + bad};
  return DGPU_OK;
}

// The Front thermal-wind coefficients of RhsArgs / AdjRhsArgs (null pointers
// when the flavour has none, forcing.f90:50).
template <typename A>
void set_front(const dgpu_handle* h, A& a) {
  const dgpu_params& p = h->p;
  if (!h->d_front.p) return;
  a.fr_yf = h->d_front.p;
  a.fr_y = h->d_front.p + h->g.ny2;
  a.fr_sx = a.fr_sz = 0.0;
  for (int n = 0; n < h->nth; ++n) {
    a.fr_sx += p.dthdx[n];
    a.fr_sz += p.dthdz[n];
  }
  a.fr_dr = p.delta / p.ro_inv;
}

// The launches of the RHS pass (k_rhs).  One launch holds every product slot
// of the substep in LDS when they fit; otherwise the outputs are split: the
// velocity (cr1, cr3 and cr2, or each on its own) and then as many scalars
// per launch as fit.  A launch copies the twiddle table to LDS when there is
// room for it beside the slots.
struct RhsPart {
  int part, th0, th1, tw_lds, ng;
  size_t lds;
};

template <typename R>
std::vector<RhsPart> rhs_plan(const dgpu_handle* h, int nth) {
  const Geo& g = h->g;
  const int G = g.nz / 8;
  const size_t L = (size_t)fft_lds_len(g.nz);
  // a launch of nslot slots: its group count and LDS size, with or without the table
  auto make = [&](int part, int th0, int th1, RhsPart& out) -> bool {
    const int nslot = ((part & 1) ? 4 : 0) + ((part & 2) ? 2 : 0) + 2 * (th1 - th0);
    int ng = 512 / G;
    if (ng < 1) ng = 1;
    if (ng > nslot) ng = nslot;
    const int rounds = (nslot + ng - 1) / ng;
    const size_t buf = (size_t)rounds * ng * L * sizeof(cpx<R>);
    const size_t tab = (size_t)g.nz * sizeof(cpx<R>);
    if (buf > h->lds_limit) return false;
    // the table goes to LDS whenever it fits: at 512 points in FP64 the slot
    // buffers and the table take 80 KB, so two workgroups still share a CU
    // (as their 128 VGPRs allow), and the pass read its twiddles from the
    // caches at 4.64 ms per step against 4.37 from LDS (512x257x512)
    const bool twl = buf + tab <= h->lds_limit;
    out = RhsPart{part, th0, th1, twl ? 1 : 0, ng, twl ? buf + tab : buf};
    return true;
  };
  std::vector<RhsPart> plan;
  RhsPart pt;
  if (make(3, 0, nth, pt)) return {pt};
  if (make(3, 0, 0, pt))
    plan.push_back(pt);
  else if (make(1, 0, 0, pt)) {
    plan.push_back(pt);
    make(2, 0, 0, pt);
    plan.push_back(pt);
  } else
    throw ApiError{DGPU_EUNSUP, "Nz too large for the spectral RHS pass"};
  for (int n0 = 0; n0 < nth;) {
    int n1 = n0 + 1;
    if (!make(0, n0, n1, pt)) throw ApiError{DGPU_EUNSUP, "Nz too large for the spectral RHS pass"};
    while (n1 < nth && make(0, n0, n1 + 1, pt)) ++n1;
    make(0, n0, n1, pt);
    plan.push_back(pt);
    n0 = n1;
  }
  return plan;
}

// ----------------------------------------------------------- the substep --
// One rk_chan_1 (channel.f90:10-917).  With R = float (the mixed form) the
// transforms, products and LES run in FP32 while the spectral state, the
// explicit spectral terms, the CN solves and the projection stay FP64; the
// RHS pass and the solves then work on increments (RHS less state, new state
// less old), so the FP64 state is never rounded to FP32 as a whole.
template <typename R>
void rk_substep_t(dgpu_handle* h, int rk_step, double time, int time_step) {
  const dgpu_params& p = h->p;
  const Geo& g = h->g;
  const int rs = rk_step - 1;
  const int nth = h->nth;
  const int Ny = g.ny;
  const bool incr = sizeof(R) < sizeof(double);
  const double temp3 = h->zeta_bar[rs] * h->h_bar[rs];  // channel.f90:29-33
  const double temp4 = h->h_bar[rs];
  const double temp5 = h->beta_bar[rs] * h->h_bar[rs];
  const int nslot = 6 + 2 * nth;
  const int ncr = 3 + nth;
  // LES gate (channel.f90:259)
  const bool les = p.use_les && (!p.create_new_flow || time_step > 100);
  // Without LES every CN matrix row depends on j alone and the wall values are
  // uniform in x and z, so the vertical solves commute with the horizontal
  // transforms: they run on the dealiased spectra (FP64 in both precision
  // forms), and cr never goes to physical space nor the solution back.
  const bool spec = h->kn.spectral_cn && !p.use_les;
  // The explicit spectral terms (channel.f90:39-253) read the state before
  // les_chan applies its wall values.  Those change planes 0, 1, Ny, Ny+1
  // only, so the LES path keeps a copy of those planes for the RHS pass.
  if (les) {
    WallCopy w;
    std::memset(&w, 0, sizeof(w));
    for (int f = 0; f < 3 + nth; ++f) {
      w.src[f] = f < 3 ? h->cu[f].p : h->cth[f - 3].p;
      w.dst[f] = h->wpre[f].p;
    }
    w.nf = 3 + nth;
    {
      Timed tm(h, DGPU_K_COPY, 32.0 * 4 * g.sp * w.nf);
      hipLaunchKernelGGL(k_copy_walls, dim3((unsigned)((g.sp + 255) / 256), 4, (unsigned)w.nf), dim3(256), 0,
                         h->stream, w, g);
      HIPCHECK(hipGetLastError());
    }
    les_stage<R>(h, time);
    if (h->les_cap) {  // the adjoint's recomputation keeps the post-BC state
      for (int f = 0; f < 3; ++f) copy_spec(h, h->cap_cub[f].p, h->cu[f].p);
      for (int n = 0; n < nth; ++n) copy_spec(h, h->cap_cub[3 + n].p, h->cth[n].p);
    }
  } else {
    // 1. state to physical space (channel.f90:279-287)
    const double2* C[3 + DGPU_MAX_TH];
    R* U[3 + DGPU_MAX_TH];
    cpx<R>* T[3 + DGPU_MAX_TH];
    for (int f = 0; f < 3 + nth; ++f) {
      C[f] = f < 3 ? h->cu[f].p : h->cth[f - 3].p;
      U[f] = f < 3 ? phys<R>(h->u[f]) : phys<R>(h->th[f - 3]);
      T[f] = xspec<R>(h->tpool[f]);
    }
    to_physical_fields<R>(h, 3 + nth, C, U, T);
  }

  // 2. products, x-transformed (channel.f90:299-459, 557-608)
  {
    ProdAllArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.base.u1 = phys<R>(h->u[0]);
    pa.base.u2 = phys<R>(h->u[1]);
    pa.base.u3 = phys<R>(h->u[2]);
    pa.base.dy = h->d_dy.p;
    pa.base.dyf = h->d_dyf.p;
    pa.base.kx = h->d_kx.p;
    pa.base.js = h->jstart;
    pa.base.je = h->jend;
    pa.base.jst = 1 << 30;
    pa.base.jet = -1;
    pa.nth = nth;
    for (int n = 0; n < nth; ++n) {
      pa.th[n] = phys<R>(h->th[n]);
      pa.jst[n] = h->jstart_th[n];
      pa.jet[n] = h->jend_th[n];
    }
    for (int s = 0; s < nslot; ++s) pa.out[s] = h->tpool[s].p;
    constexpr int KT = XTile<R>::KT;
    // a row's group needs Nx/8 lanes whatever the limit
    const size_t kxb = g.nq * sizeof(R);  // kx behind the table
    const int ngx = x_ngroups<R>(h, KT, std::max(sizeof(R) == 8 ? h->kn.xprod_threads : h->kn.xprod_threads32, g.nx / 8),
                                 kxb);
    const size_t lds = x_lds<R>(h, ngx) + kxb;
    if (lds > h->lds_limit) throw ApiError{DGPU_EUNSUP, "Nx too large for the products pass"};
    dim3 grid((g.nz + KT - 1) / KT, Ny), block(ngx * (g.nx / 8));
    // reads u1, u2, u3, th once (neighbour planes from cache); writes the product x-spectra
    Timed tm(h, DGPU_K_XPROD, (sizeof(R) * g.pp * (3 + nth) + 2.0 * sizeof(R) * g.tp * nslot) * Ny,
             2.0 * sizeof(R) * g.tp * nslot * Ny);
    // 512-point FP64 rows (the benchmark's grid) take the compile-time
    // transform: 5.77 -> 5.54 ms per step at 512x257x512 (the FP32 pass was
    // 3% slower with it, 2.78 -> 2.87 ms, and keeps the run-time one)
    auto kern = g.nx <= 512 ? k_xprod_all<R, true> : k_xprod_all<R, false>;
    if constexpr (sizeof(R) == 8)
      if (g.nx == 512 && (h->kn.fft_const & 1)) kern = k_xprod_all<R, true, 9>;
    if (g.nz % KT) kern = g.nx <= 512 ? k_xprod_all<R, true, 0, true> : k_xprod_all<R, false, 0, true>;
    // FP64 loads the next input phase behind each pair (5.53 -> 5.37 ms per
    // step at 512x257x512, still two waves per SIMD at 226 VGPRs).  With the
    // Stockham transform the FP32 pass keeps the phase-start loads: its 145
    // VGPRs with them cost a wave per SIMD (4 -> 3), 2.85 -> 3.30 ms; with the
    // cross-lane transform it fits 128 VGPRs and gains (below)
    if constexpr (sizeof(R) == 8)
      if (h->kn.xprod_pf & 1) {
        kern = g.nx <= 512 ? k_xprod_all<R, true, 0, false, true> : k_xprod_all<R, false, 0, false, true>;
        if (g.nx == 512 && (h->kn.fft_const & 1)) kern = k_xprod_all<R, true, 9, false, true>;
        if (g.nz % KT)
          kern = g.nx <= 512 ? k_xprod_all<R, true, 0, true, true> : k_xprod_all<R, false, 0, true, true>;
      }
    if (g.nx == 512 && g.nz % KT == 0 && use_xl<R>(h, XL_PROD)) {
      if constexpr (sizeof(R) == 8) {
        const bool pf = h->kn.xprod_pf & 1;
        if (xl_powers<R>(h, XL_PROD))
          kern = pf ? k_xprod_all<R, true, FFT_XL512P, false, true> : k_xprod_all<R, true, FFT_XL512P>;
        else
          kern = pf ? k_xprod_all<R, true, FFT_XL512, false, true> : k_xprod_all<R, true, FFT_XL512>;
      }
      else  // 128 VGPRs with the early loads, four waves per SIMD: 2.61 -> 2.32 ms per step
        kern = (h->kn.xprod_pf & 2) ? k_xprod_all<R, true, FFT_XL512, false, true> : k_xprod_all<R, true, FFT_XL512>;
    }
    hipLaunchKernelGGL(kern, grid, block, lds, h->stream, pa, g, twx_of<R>(h), 1);
    HIPCHECK(hipGetLastError());
  }

  // 3. explicit spectral terms, cf, cr, backward z transform of cr
  {
    RhsArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int s = 0; s < nslot; ++s) a.X[s] = h->tpool[s].p;
    for (int c = 0; c < ncr; ++c) a.Tcr[c] = h->tpool[nslot + c].p;
    a.cu1 = h->cu[0].p;
    a.cu2 = h->cu[1].p;
    a.cu3 = h->cu[2].p;
    if (les) {
      a.wcu1 = h->wpre[0].p;
      a.wcu2 = h->wpre[1].p;
      a.wcu3 = h->wpre[2].p;
      for (int n = 0; n < nth; ++n) a.wcth[n] = h->wpre[3 + n].p;
    }
    a.cp = h->cp.p;
    a.incr = (!incr || spec) ? 0 : (les ? 1 : 2);
    if (spec)  // cr in the (otherwise idle) cr x-spectra buffers, at plane pitch sp
      for (int c = 0; c < ncr; ++c) a.crs[c] = h->tpool[nslot + c].p;
    if (spec && incr) {  // the mixed form: cr - cu as FP32 spectra
      a.incr = 2;
      a.crs32 = 1;
    }
    a.cub1 = h->cu[0].p;
    a.cub2 = h->cu[1].p;
    a.cub3 = h->cu[2].p;
    if (les) {
      a.lesf1 = h->lesf[0].p;
      a.lesf2 = h->lesf[1].p;
      a.lesf3 = h->lesf[2].p;
      for (int n = 0; n < nth; ++n) a.lesfth[n] = h->lesfth[n].p;
    }
    a.cf1 = h->cf[0].p;
    a.cf2 = h->cf[1].p;
    a.cf3 = h->cf[2].p;
    for (int n = 0; n < nth; ++n) {
      a.cth[n] = h->cth[n].p;
      a.cthb[n] = h->cth[n].p;
      a.cfth[n] = h->cfth[n].p;
      a.pr[n] = p.pr[n];
      a.dthdx[n] = p.dthdx[n];
      a.dthdz[n] = p.dthdz[n];
      a.jst[n] = h->jstart_th[n];
      a.jet[n] = h->jend_th[n];
    }
    a.kx = h->d_kx.p;
    a.kz = h->d_kz.p;
    a.dy = h->d_dy.p;
    a.dyf = h->d_dyf.p;
    a.nu = h->nu;
    a.beta = p.beta;
    a.coriolis = p.ro_inv / p.delta;
    a.grav_x = p.grav_x;
    a.grav_y = p.grav_y;
    a.grav_z = p.grav_z;
    a.temp3 = temp3;
    a.temp4 = temp4;
    a.temp5 = temp5;
    if (p.f_type == 1)
      a.force = p.px0;  // channel.f90:108-112
    else if (p.f_type == 2)
      a.force = p.px0 + p.amp_omega0 * std::cos(p.omega0 * time);  // :113-118
    else
      a.force = 0.0;
    set_front(h, a);
    a.rk_step = rk_step;
    a.nth = nth;
    a.js = h->jstart;
    a.je = h->jend;
    auto kern = g.nz <= 512 ? k_rhs<R, true> : k_rhs<R, false>;
    if (g.nz == 512 && (h->kn.fft_const & (sizeof(R) == 8 ? 1 : 2))) kern = k_rhs<R, true, 9>;
    if (g.nz == 512 && use_xl<R>(h, XL_RHS)) kern = k_rhs<R, true, FFT_XL512>;
    // x-spectra of the products in, cr x-spectra out; cu1..3, cp, cth, cf1..3,
    // cfth read and cf1..3, cfth written (neighbour planes come from cache)
    const double crbytes = spec ? (incr ? 8.0 : 16.0) * g.sp * ncr : 2.0 * sizeof(R) * g.tp * ncr;
    Timed tm(h, DGPU_K_RHS,
             (2.0 * sizeof(R) * g.tp * nslot + crbytes + 16.0 * g.sp * (4 + 2 * nth) +
              2.0 * sizeof(R) * g.sp * 2 * (3 + nth)) * Ny,  // the carry cf is stored in R
             (2.0 * sizeof(R) * g.tp * nslot + (spec ? 0.0 : crbytes)) * Ny);
    if (h->zgen) {
      // general Nz: the z spectra of the slots first (planes 1..Ny), then a
      // combine-only pass; without spectral solves cr goes to the slot
      // buffers (each cr slot overwrites a slot its own outputs read first,
      // mode by mode in the same thread) and through the general z inverse
      for (int s0 = 0; s0 < nslot; s0 += MAXBF) {
        FieldBatch fb{};
        fb.nf = std::min(MAXBF, nslot - s0);
        for (int k = 0; k < fb.nf; ++k) {
          fb.in[k] = h->tpool[s0 + k].p;
          fb.out[k] = h->zpool[s0 + k].p;
        }
        zgen_fwd<R>(h, xspec<R>(h->tpool[s0]), h->zpool[s0].p, 1, Ny, false, 1.0, 1.0, &fb);
      }
      for (int s = 0; s < nslot; ++s) a.X[s] = h->zpool[s].p;
      if (!spec)
        for (int c = 0; c < ncr; ++c) a.crs[c] = h->zpool[c].p;
      a.part = 3;
      a.th0 = 0;
      a.th1 = nth;
      hipLaunchKernelGGL((k_rhs<R, false, 0, true>), dim3(g.nq, Ny), dim3(64), 0, h->stream, a, g, twz_of<R>(h),
                         1);
      HIPCHECK(hipGetLastError());
      if (!spec)
        for (int c0 = 0; c0 < ncr; c0 += MAXBF) {
          FieldBatch fb{};
          fb.nf = std::min(MAXBF, ncr - c0);
          for (int k = 0; k < fb.nf; ++k) {
            fb.in[k] = h->zpool[c0 + k].p;
            fb.out[k] = h->tpool[nslot + c0 + k].p;
          }
          zgen_inv<R>(h, h->zpool[c0].p, xspec<R>(h->tpool[nslot + c0]), 1, Ny, OP_ID, 1.0, 1.0, &fb);
        }
    } else {
      for (const RhsPart& pt : rhs_plan<R>(h, nth)) {
        a.part = pt.part;
        a.th0 = pt.th0;
        a.th1 = pt.th1;
        a.tw_lds = pt.tw_lds;
        dim3 grid(g.nq, Ny), block(pt.ng * (g.nz / 8));
        hipLaunchKernelGGL(kern, grid, block, pt.lds, h->stream, a, g, twz_of<R>(h), 1);
        HIPCHECK(hipGetLastError());
      }
    }
  }

  // 4. cr to physical space (second half of fft_xz_to_physical, :485-487, :621)
  // slot order of cr: 0 cr1, 1 cr3, 2 cr2, 3+n crth_n
  if (!spec) {
    cpx<R>* TC[3 + DGPU_MAX_TH];
    R* RC[3 + DGPU_MAX_TH];
    for (int c = 0; c < 3 + nth; ++c) {
      TC[c] = xspec<R>(h->tpool[nslot + c]);
      RC[c] = c == 0 ? phys<R>(h->r[0]) : c == 1 ? phys<R>(h->r[2]) : c == 2 ? phys<R>(h->r[1]) : phys<R>(h->rth[c - 3]);
    }
    xinv_fields<R>(h, 3 + nth, TC, RC, 1, Ny);
  }

  // 5. implicit vertical solves (channel.f90:656-877)
  double transient = h->w_transient;
  if (p.w_bc_ymax == 1 && p.f_type == 4) {
    transient = p.w_bc_ymax_c1 + p.amp_omega0 * std::sin(p.omega0 * (p.ro_inv / p.delta * time - p.force_start));
    h->w_transient = transient;
  }
  if (p.use_les) {
    // with use_LES the eddy viscosity enters every CN row (channel.f90:519-547, 636-689);
    // every field's solve in one launch
    LesCnFields<R> mf{};
    auto add = [&](int field, DevBuf<double>& rb, DevBuf<double>& ub, const double* nt, int ca, int cb) {
      const int k = mf.nf++;
      mf.r[k] = phys<R>(rb);
      mf.u[k] = phys<R>(ub);
      mf.nt[k] = nt;
      mf.rows[k] = h->trirows.p + ((size_t)field * 3 + rs) * g.ny2;
      mf.ca[k] = ca;
      mf.cb[k] = cb;
    };
    for (int n = 0; n < nth; ++n)
      add(3 + n, h->rth[n], h->th[n], h->kappa_t[n].p, h->jstart_th[n] - 1, h->jend_th[n] + 1);
    add(1, h->r[1], h->u[1], h->nu_t.p, 1, Ny + 1);
    add(0, h->r[0], h->u[0], h->nu_t.p, h->jstart - 1, h->jend + 1);
    add(2, h->r[2], h->u[2], h->nu_t.p, h->jstart - 1, h->jend + 1);
    launch_thomas_les<R>(h, 0, rs, mf.r[0], mf.u[0], mf.nt[0], mf.ca[0], mf.cb[0], transient, incr, &mf);
  } else if (spec) {
    // every field's solve in one launch (cr slot order: 0 cr1, 1 cr3, 2 cr2, 3+n crth_n)
    CnFields<double> mf{};
    auto add = [&](int field, int slot, DevBuf<double2>& state, int ca, int cb) {
      const int k = mf.nf++;
      mf.r[k] = reinterpret_cast<const double*>(h->tpool[nslot + slot].p);
      mf.u[k] = reinterpret_cast<double*>(state.p);
      mf.rows[k] = h->rows.p + ((size_t)field * 3 + rs) * g.ny2;
      mf.ca[k] = ca;
      mf.cb[k] = cb;
    };
    add(0, 0, h->cu[0], h->jstart - 1, h->jend + 1);
    add(1, 2, h->cu[1], 1, Ny + 1);
    add(2, 1, h->cu[2], h->jstart - 1, h->jend + 1);
    for (int n = 0; n < nth; ++n) add(3 + n, 3 + n, h->cth[n], h->jstart_th[n] - 1, h->jend_th[n] + 1);
    launch_thomas<double>(h, 0, rs, static_cast<const double*>(mf.r[0]), mf.u[0], mf.ca[0], mf.cb[0], transient,
                          false, nullptr, true, &mf, incr);
    for (int k = 0; k < mf.nf; ++k) launch_herm0(h, reinterpret_cast<double2*>(mf.u[k]), mf.ca[k], mf.cb[k]);
  } else {
    for (int n = 0; n < nth; ++n)
      launch_thomas<R>(h, 3 + n, rs, phys<R>(h->rth[n]), phys<R>(h->th[n]), h->jstart_th[n] - 1,
                       h->jend_th[n] + 1, transient, incr);
    launch_thomas<R>(h, 1, rs, phys<R>(h->r[1]), phys<R>(h->u[1]), 1, Ny + 1, transient, incr);
    launch_thomas<R>(h, 0, rs, phys<R>(h->r[0]), phys<R>(h->u[0]), h->jstart - 1, h->jend + 1, transient, incr);
    launch_thomas<R>(h, 2, rs, phys<R>(h->r[2]), phys<R>(h->u[2]), h->jstart - 1, h->jend + 1, transient, incr);
  }

  if (h->les_cap) {  // ... and the CN solutions
    const size_t np = (size_t)h->g.ny2 * h->g.pp;
    for (int f = 0; f < 3; ++f)
      HIPCHECK(hipMemcpyAsync(h->cap_x[f].p, h->u[f].p, sizeof(double) * np, hipMemcpyDeviceToDevice, h->stream));
    for (int n = 0; n < nth; ++n)
      HIPCHECK(hipMemcpyAsync(h->cap_x[3 + n].p, h->th[n].p, sizeof(double) * np, hipMemcpyDeviceToDevice,
                              h->stream));
  }

  // 6. back to Fourier space on the solved planes (channel.f90:891-896); the
  // mixed form adds the transformed increment to the FP64 state
  if (!spec) {
    // fields that share their plane range go in one launch per pass
    struct Sol {
      R* u;
      double2* c;
      int j0, j1;
    };
    std::vector<Sol> sol;
    sol.push_back({phys<R>(h->u[0]), h->cu[0].p, h->jstart - 1, h->jend + 1});
    sol.push_back({phys<R>(h->u[1]), h->cu[1].p, 1, Ny + 1});
    sol.push_back({phys<R>(h->u[2]), h->cu[2].p, h->jstart - 1, h->jend + 1});
    for (int n = 0; n < nth; ++n) sol.push_back({phys<R>(h->th[n]), h->cth[n].p, h->jstart_th[n] - 1, h->jend_th[n] + 1});
    std::vector<bool> done(sol.size(), false);
    for (size_t a = 0; a < sol.size(); ++a) {
      if (done[a]) continue;
      R* U[3 + DGPU_MAX_TH];
      double2* CC[3 + DGPU_MAX_TH];
      cpx<R>* TT[3 + DGPU_MAX_TH];
      int k = 0;
      for (size_t b = a; b < sol.size(); ++b)
        if (!done[b] && sol[b].j0 == sol[a].j0 && sol[b].j1 == sol[a].j1) {
          done[b] = true;
          U[k] = sol[b].u;
          CC[k] = sol[b].c;
          TT[k] = xspec<R>(h->tpool[b]);
          ++k;
        }
      to_fourier_fields<R>(h, k, U, CC, TT, sol[a].j0, sol[a].j1, incr);
    }
  }

  // courant on the last substep of every update_dt-th step, on the velocity
  // just solved (channel.f90:882-888); the new delta_t applies from the next
  // substep, after this one has used temp4
  const bool cfl_now = h->var_dt && rk_step == 3 && time_step % h->cour.update_dt == 0;
  double cfl_pts = 0.0;
  if (cfl_now) {
    if (incr || spec) {  // u holds increments or the substep's input: take the new FP64 state
      for (int f = 0; f < 3; ++f) to_physical<double>(h, h->cu[f].p, h->r[f].p, xspec<double>(h->tpool[f]));
      cfl_pts = cfl_min(h, h->r[0].p, h->r[1].p, h->r[2].p, false);
    } else {
      cfl_pts = cfl_min(h, h->u[0].p, h->u[1].p, h->u[2].p, false);
    }
  }

  // 7. projection and pressure (channel.f90:901-911)
  launch_remdiv(h, true, temp4);
  if (cfl_now) set_delta_t_impl(h, courant_dt(h, cfl_pts, time));
}

void rk_substep(dgpu_handle* h, int rk_step, double time, int time_step) {
  if (h->precision == DGPU_MIXED)
    rk_substep_t<float>(h, rk_step, time, time_step);
  else
    rk_substep_t<double>(h, rk_step, time, time_step);
}

// ------------------------------------------------------------- adjoint --
// Fields of one substep: 0 u1, 1 u2, 2 u3, 3+n th_n; their solved planes.
void solve_range(const dgpu_handle* h, int f, int* ca, int* cb) {
  if (f == 1) {
    *ca = 1;
    *cb = h->g.ny + 1;
  } else if (f < 3) {
    *ca = h->jstart - 1;
    *cb = h->jend + 1;
  } else {
    *ca = h->jstart_th[f - 3] - 1;
    *cb = h->jend_th[f - 3] + 1;
  }
}

// Transposed CN matrices of the adjoint, all rows plain:
// (A^T) sub(j) = c(j-1), sup(j) = a(j+1).
void build_adj_tables(dgpu_handle* h) {
  const int n = h->g.ny2;
  std::vector<SegRow> rt((size_t)h->nfields * 3 * n);
  for (int f = 0; f < h->nfields; ++f)
    for (int rs = 0; rs < 3; ++rs) {
      Tri t = build_tri(h, f, rs);
      Tri tt = t;
      for (int j = 0; j < n; ++j) {
        tt.a[j] = (j >= 1) ? t.c[j - 1] : 0.0;
        tt.c[j] = (j + 1 < n) ? t.a[j + 1] : 0.0;
        tt.kind[j] = 3;
        tt.wp[j] = tt.wm[j] = tt.vb[j] = 0.0;
      }
      factor_rows(tt, h->seg_L, rt.data() + ((size_t)f * 3 + rs) * n);
    }
  upload(h->rowsT, rt);
}

void adj_alloc(dgpu_handle* h) {
  if (h->adj_ready) return;
  const Geo& g = h->g;
  const int nth = h->nth, nf = 3 + nth, np_ = 4 + 2 * nth;
  const size_t ns = (size_t)g.ny2 * g.sp, npx = (size_t)g.ny2 * g.pp;
  auto zs = [&](DevBuf<double2>& b) {
    b.alloc(ns);
    HIPCHECK(hipMemset(b.p, 0, sizeof(double2) * ns));
  };
  auto zp = [&](DevBuf<double>& b) {
    b.alloc(npx);
    HIPCHECK(hipMemset(b.p, 0, sizeof(double) * npx));
  };
  for (int f = 0; f < 3; ++f) {
    zs(h->a_cu[f]);
    zs(h->a_cf[f]);
  }
  zs(h->a_cp);
  for (int n = 0; n < nth; ++n) {
    zs(h->a_cth[n]);
    zs(h->a_cfth[n]);
  }
  for (int f = 0; f < nf; ++f) {
    zs(h->a_cr[f]);
    zs(h->a_fb[f]);
    zp(h->a_rb[f]);
    zp(h->a_ub[f]);
  }
  for (int q = 0; q < np_; ++q) {  // planes 0 and Ny+1 stay zero
    zp(h->a_pA[q]);
    zp(h->a_pB[q]);
  }
  build_adj_tables(h);
  h->d_w.alloc(g.ny2);
  h->adj_ready = true;
}

double2* adj_spec(dgpu_handle* h, int f) { return f < 3 ? h->a_cu[f].p : h->a_cth[f - 3].p; }
const double2* fwd_spec(dgpu_handle* h, int f) { return f < 3 ? h->cu[f].p : h->cth[f - 3].p; }

// The vector-Jacobian product of rk_chan_1 (FP64 DNS): on entry the forward
// state is the substep's input and the adjoint state is the gradient of the
// loss with respect to the substep's output; on exit the adjoint state is the
// gradient with respect to the input.  The forward state is not changed.
// ------------------------------------------------- adjoint: LES terms --
void les_adj_alloc(dgpu_handle* h) {
  if (h->les_adj_ready) return;
  const Geo& g = h->g;
  const int nth = h->nth, nf = 3 + nth;
  const size_t ns = (size_t)g.ny2 * g.sp, npx = (size_t)g.ny2 * g.pp;
  for (int f = 0; f < nf; ++f) {
    h->cap_cub[f].alloc(ns);
    h->cap_x[f].alloc(npx);
    h->a_cub[f].alloc(ns);
  }
  h->fsave.resize(7 + 2 * nth);
  for (auto& b : h->fsave) b.alloc(ns);
  h->a_nt.alloc(npx);
  for (int n = 0; n < nth; ++n) {
    h->a_kt[n].alloc(npx);
    for (int w = 0; w < 3; ++w) h->a_T[w][n].alloc(npx);
  }
  for (auto& b : h->a_G) b.alloc(npx);
  h->les_adj_ready = true;
}

// The spectral state buffers in the order of fsave.
std::vector<double2*> state_bufs(dgpu_handle* h) {
  std::vector<double2*> v = {h->cu[0].p, h->cu[1].p, h->cu[2].p, h->cp.p, h->cf[0].p, h->cf[1].p, h->cf[2].p};
  for (int n = 0; n < h->nth; ++n) {
    v.push_back(h->cth[n].p);
    v.push_back(h->cfth[n].p);
  }
  return v;
}

// Runs the forward substep from the current state, keeping what the adjoint
// of the LES terms reads: the post-BC state, the CN solutions, and the
// gradients, nu_t and kappa_t of the LES stage; then puts the input state
// back and rebuilds the physical post-BC fields in u, th.
void les_adj_recompute(dgpu_handle* h, int rk_step, double time, int time_step) {
  les_adj_alloc(h);
  std::vector<double2*> st = state_bufs(h);
  for (size_t k = 0; k < st.size(); ++k) copy_spec(h, h->fsave[k].p, st[k]);
  const double wt = h->w_transient;
  h->les_cap = true;
  try {
    rk_substep_t<double>(h, rk_step, time, time_step);
  } catch (...) {
    h->les_cap = false;
    throw;
  }
  h->les_cap = false;
  for (size_t k = 0; k < st.size(); ++k) copy_spec(h, st[k], h->fsave[k].p);
  h->w_transient = wt;
  for (int f = 0; f < 3; ++f) to_physical<double>(h, h->cap_cub[f].p, h->u[f].p, xspec<double>(h->tpool[0]));
  for (int n = 0; n < h->nth; ++n)
    to_physical<double>(h, h->cap_cub[3 + n].p, h->th[n].p, xspec<double>(h->tpool[0]));
}

// Adjoint of les_stage (les.f90:20-1455 as arranged in les_stage): from the
// adjoints of the subgrid forcings (the fbar of the RHS pass) and of nu_t,
// kappa_t (from the CN solves) to the post-BC state adjoint a_cub and the
// physical state adjoints a_ub (u1, u3 enter the model directly).
void les_adj_stage(dgpu_handle* h, int rk_step) {
  (void)rk_step;
  const dgpu_params& p = h->p;
  const Geo& g = h->g;
  const int nth = h->nth, Ny = g.ny, n2 = g.ny2;
  const double NxNz = (double)g.nx * g.nz;
  const size_t npx = (size_t)n2 * g.pp;
  cpx<double>* T0 = xspec<double>(h->tpool[0]);
  auto grid1 = [](long long n) { return dim3((unsigned)((n + 255) / 256)); };
  // 1. transposed stress divergence
  {
    AdjLesDivArgs a;
    std::memset(&a, 0, sizeof(a));
    a.f1 = h->a_fb[0].p;
    a.f2 = h->a_fb[1].p;
    a.f3 = h->a_fb[2].p;
    a.t11 = h->ctau[0].p;
    a.t33 = h->ctau[1].p;
    a.t13 = h->ctau[2].p;
    a.t12 = h->ctau[3].p;
    a.t23 = h->ctau[4].p;
    a.c1 = h->ctau[5].p;
    a.c3 = h->ctau[6].p;
    for (int n = 0; n < nth; ++n) {
      a.fth[n] = h->a_fb[3 + n].p;
      a.th_f1[n] = h->cthf[0][n].p;
      a.th_f3[n] = h->cthf[1][n].p;
      a.jst[n] = h->jstart_th[n];
      a.jet[n] = h->jend_th[n];
    }
    a.kx = h->d_kx.p;
    a.kz = h->d_kz.p;
    a.dyf = h->d_dyf.p;
    a.js = h->jstart;
    a.j2 = Ny - 1;
    a.nth = nth;
    a.th_sign = (p.les_model_type == 5) ? -1 : 1;
    Timed tm(h, DGPU_K_ADJOINT, 16.0 * (3 + 7 + 3 * nth) * g.sp * n2);
    hipLaunchKernelGGL(k_adj_les_div, grid1(g.sp * n2), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  // 2. F^T of the stress and flux spectra (the forward transformed planes 1..Ny)
  auto ftrans = [&](const double2* c, double* out) {
    HIPCHECK(hipMemsetAsync(out, 0, sizeof(double) * npx, h->stream));
    launch_zinv<double>(h, c, T0, 1, Ny, 1.0 / NxNz, 0.5 / NxNz);
    launch_xinv<double>(h, T0, out, 1, Ny);
  };
  for (int m = 0; m < 7; ++m) ftrans(h->ctau[m].p, h->tau[m].p);
  for (int n = 0; n < nth; ++n)
    for (int w = 0; w < 2; ++w) ftrans(h->cthf[w][n].p, h->thf[w][n].p);
  // 3. the scalar model, then the velocity model
  LesVelArgs<double> fv;
  std::memset(&fv, 0, sizeof(fv));
  for (int m = 0; m < 9; ++m) fv.G[m] = h->grad[m].p;
  fv.u1 = h->u[0].p;
  fv.u3 = h->u[2].p;
  fv.dy = h->d_dy.p;
  fv.dyf = h->d_dyf.p;
  fv.dx1 = p.lx / p.nx;
  fv.dz1 = p.lz / p.nz;
  fv.type = p.les_model_type;
  fv.js = h->jstart;
  fv.je = h->jend;
  const long long npt = (long long)npx;
  for (int n = 0; n < nth; ++n) {
    AdjLesThArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int m = 0; m < 9; ++m) a.fw.G[m] = h->grad[m].p;
    a.fw.T1 = h->dth[0][n].p;
    a.fw.T2 = (p.les_model_type == 5) ? h->dth[1][n].p : h->dth[0][n].p;
    a.fw.T3 = h->dth[2][n].p;
    a.fw.nu_t = h->nu_t.p;
    a.fw.dy = h->d_dy.p;
    a.fw.dyf = h->d_dyf.p;
    a.fw.dx1 = p.lx / p.nx;
    a.fw.dz1 = p.lz / p.nz;
    a.fw.type = p.les_model_type;
    a.fw.jst = h->jstart_th[n];
    a.fw.jet = h->jend_th[n];
    a.f1bar = h->thf[0][n].p;
    a.f3bar = h->thf[1][n].p;
    a.ktbar = h->a_kt[n].p;
    for (int m = 0; m < 9; ++m) a.gbar[m] = h->a_G[m].p;
    a.gbar_add = (n > 0) ? 1 : 0;
    for (int w = 0; w < 3; ++w) a.tbar[w] = h->a_T[w][n].p;
    a.ntbar = h->a_nt.p;
    Timed tm(h, DGPU_K_ADJOINT, 8.0 * npt * 40);
    hipLaunchKernelGGL(k_adj_les_th, grid1(npt), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  {
    AdjLesVelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.fw = fv;
    const int bmap[8] = {0, 1, 2, 3, 4, -1, 5, 6};  // tau order t11 t33 t13 t12 t23 cr1 cr3
    for (int k = 0; k < 8; ++k) a.bar[k] = (bmap[k] < 0) ? h->a_nt.p : h->tau[bmap[k]].p;
    const bool th_grads = (p.les_model_type == 5 && nth > 0);
    for (int m = 0; m < 9; ++m) {
      a.gbar_in[m] = th_grads ? h->a_G[m].p : nullptr;
      a.gbar[m] = h->a_G[m].p;
    }
    a.u1bar = h->a_ub[0].p;
    a.u3bar = h->a_ub[2].p;
    Timed tm(h, DGPU_K_ADJOINT, 8.0 * npt * 60);
    hipLaunchKernelGGL(k_adj_les_vel, grid1(npt), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  // 4. gradients back to the post-BC spectral state: P^T, then D^T
  double2* X = h->a_cr[0].p;
  auto gback = [&](const double* gb, int op, double2* target) {
    launch_xfwd<double>(h, gb, T0, 0, n2 - 1);
    launch_zfwd<double>(h, T0, X, 0, n2 - 1, false, NxNz, 2.0 * NxNz);
    Timed tm(h, DGPU_K_ADJOINT, 48.0 * g.sp * n2);
    hipLaunchKernelGGL(k_adj_deriv, grid1(g.sp * n2), dim3(256), 0, h->stream, X, target, op, h->d_kx.p,
                       h->d_kz.p, h->d_dy.p, h->d_dyf.p, g);
    HIPCHECK(hipGetLastError());
  };
  const int gsrc[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
  const int gop[9] = {OP_DX, OP_DY_GY, OP_DZ, OP_DX, OP_DY_GYF, OP_DZ, OP_DX, OP_DY_GY, OP_DZ};
  for (int m = 0; m < 9; ++m) gback(h->a_G[m].p, gop[m], h->a_cub[gsrc[m]].p);
  for (int n = 0; n < nth; ++n) {
    const int ops[3] = {OP_DX, OP_DY_GY, OP_DZ};
    for (int w = 0; w < 3; ++w) {
      if (w == 1 && p.les_model_type != 5) continue;
      gback(h->a_T[w][n].p, ops[w], h->a_cub[3 + n].p);
    }
  }
}

void adj_substep(dgpu_handle* h, int rk_step, double time, int time_step) {
  const dgpu_params& p = h->p;
  const Geo& g = h->g;
  if (h->precision != DGPU_FP64) throw ApiError{DGPU_EUNSUP, "the adjoint covers the FP64 substep"};
  if (h->var_dt) throw ApiError{DGPU_EUNSUP, "the adjoint covers a fixed time step (variable_dt is on)"};
  adj_alloc(h);
  // the LES gate of the forward substep (channel.f90:259); with it closed
  // nu_t and kappa_t stay zero and the substep is the DNS one
  const bool les = p.use_les && (!p.create_new_flow || time_step > 100);
  if (les) les_adj_recompute(h, rk_step, time, time_step);
  const int rs = rk_step - 1;
  const int nth = h->nth, nf = 3 + nth, Ny = g.ny, n2 = g.ny2;
  const double temp3 = h->zeta_bar[rs] * h->h_bar[rs];
  const double temp4 = h->h_bar[rs];
  const double temp5 = h->beta_bar[rs] * h->h_bar[rs];
  const double NxNz = (double)g.nx * g.nz;
  const size_t ns = (size_t)n2 * g.sp;
  const size_t npx = (size_t)n2 * g.pp;
  cpx<double>* T0 = xspec<double>(h->tpool[0]);
  auto grid1 = [](long long n) { return dim3((unsigned)((n + 255) / 256)); };

  // 1. projection and pressure update (channel.f90:901-911, 946-1023)
  {
    AdjRemDivArgs a;
    a.a1 = h->a_cu[0].p;
    a.a2 = h->a_cu[1].p;
    a.a3 = h->a_cu[2].p;
    a.ap = h->a_cp.p;
    a.ib = h->rd_ib.p;
    a.kx = h->d_kx.p;
    a.kz = h->d_kz.p;
    a.inv_h = 1.0 / temp4;
    a.js = h->jstart;
    a.je = h->jend;
    Timed tm(h, DGPU_K_ADJOINT, 128.0 * g.sp * Ny);
    // segments and workgroup shape as in launch_remdiv
    const int nr = g.ny - 1;
    auto launch = [&](auto kern, int L, int MW) {
      const int spw = 64 / MW;
      const int S = ((nr + L - 1) / L + spw - 1) / spw * spw;
      hipLaunchKernelGGL(kern, dim3((unsigned)((g.sp + MW - 1) / MW)), dim3(MW * S), 0, h->stream, a, g,
                         h->d_prow.p, S);
    };
    if (nr <= 4 * 64)
      launch(k_adj_remdiv_seg<4, 16, 1024>, 4, 16);
    else if (nr <= 8 * 64)
      launch(k_adj_remdiv_seg<8, 8, 512>, 8, 8);
    else
      launch(k_adj_remdiv_seg<8, 4, 512>, 8, 4);  // launch_remdiv has rejected ny beyond this
    HIPCHECK(hipGetLastError());
  }
  // Without LES the forward solved on the spectra (spectral_cn): so does the
  // transposed solve, in place on the state adjoint, writing crbar directly
  const bool spec = h->kn.spectral_cn && !p.use_les;
  // 2. the forward transform of the solved planes: xbar = F^T cubar there, and
  //    those planes of cubar are consumed (the other planes pass through)
  for (int f = 0; f < nf && !spec; ++f) {
    int ca, cb;
    solve_range(h, f, &ca, &cb);
    double* xb = (f < 3) ? h->r[f].p : h->rth[f - 3].p;
    HIPCHECK(hipMemsetAsync(xb, 0, sizeof(double) * npx, h->stream));
    launch_zinv<double>(h, adj_spec(h, f), T0, ca, cb, 1.0 / NxNz, 0.5 / NxNz);
    launch_xinv<double>(h, T0, xb, ca, cb);
    HIPCHECK(hipMemsetAsync(adj_spec(h, f) + (size_t)ca * g.sp, 0, sizeof(double2) * g.sp * (cb - ca + 1),
                            h->stream));
  }
  // With LES the state that the products, the LES stage and the unsolved
  // planes see is the post-BC one: its adjoint gathers in a_cub, and the
  // explicit spectral terms (which read the pre-BC state) add to a_cu.
  if (les) {
    for (int f = 0; f < nf; ++f) {
      copy_spec(h, h->a_cub[f].p, adj_spec(h, f));
      HIPCHECK(hipMemsetAsync(adj_spec(h, f), 0, sizeof(double2) * ns, h->stream));
    }
  }
  // 3. CN solves: vbar = A^-T xbar, and from it rbar and ubar = E^T vbar
  if (les) {
    // with nu_t, kappa_t in the rows (k_thomas_les): lambda, rbar, ubar_old
    // and the eddy viscosity/diffusivity adjoints
    HIPCHECK(hipMemsetAsync(h->a_nt.p, 0, sizeof(double) * npx, h->stream));
    for (int n = 0; n < nth; ++n) HIPCHECK(hipMemsetAsync(h->a_kt[n].p, 0, sizeof(double) * npx, h->stream));
    for (int f = 0; f < nf; ++f) {
      const double* xb = (f < 3) ? h->r[f].p : h->rth[f - 3].p;
      const double* uo = (f < 3) ? h->u[f].p : h->th[f - 3].p;
      const double* nt = (f < 3) ? h->nu_t.p : h->kappa_t[f - 3].p;
      double* ntb = (f < 3) ? h->a_nt.p : h->a_kt[f - 3].p;
      const TriRow* rows = h->trirows.p + ((size_t)f * 3 + rs) * n2;
      Timed tm(h, DGPU_K_ADJOINT, 96.0 * npx);
      hipLaunchKernelGGL(k_adj_thomas_les, grid1(g.pp), dim3(256), 0, h->stream, xb, h->cap_x[f].p, uo, nt,
                         h->a_rb[f].p, h->a_ub[f].p, ntb, h->sbl.p, h->sgl.p, rows, n2, (long long)g.pp);
      HIPCHECK(hipGetLastError());
    }
  }
  for (int f = 0; f < nf && !les; ++f) {
    if (spec) {
      int ca, cb;
      solve_range(h, f, &ca, &cb);
      launch_herm0(h, adj_spec(h, f), ca, cb);
      double* lam = reinterpret_cast<double*>(adj_spec(h, f));
      launch_thomas_adj(h, f, rs, lam, reinterpret_cast<double*>(h->a_cr[f].p), lam, true, ca, cb);
    } else {
      const double* xb = (f < 3) ? h->r[f].p : h->rth[f - 3].p;
      launch_thomas_adj(h, f, rs, xb, h->a_rb[f].p, h->a_ub[f].p);
    }
  }
  // 4. crbar = P^T rbar on planes 1..Ny (the inverse transform of cr)
  for (int f = 0; f < nf && !spec; ++f) {
    launch_xfwd<double>(h, h->a_rb[f].p, T0, 1, Ny);
    launch_zfwd<double>(h, T0, h->a_cr[f].p, 1, Ny, false, NxNz, 2.0 * NxNz);
  }
  // 5. the explicit spectral terms
  {
    AdjRhsArgs a;
    std::memset(&a, 0, sizeof(a));
    // natural order u1, u2, u3, th_n
    for (int f = 0; f < nf; ++f) {
      a.acr[f] = h->a_cr[f].p;
      a.fbar[f] = h->a_fb[f].p;
    }
    a.acf[0] = h->a_cf[0].p;
    a.acf[1] = h->a_cf[1].p;
    a.acf[2] = h->a_cf[2].p;
    for (int n = 0; n < nth; ++n) {
      a.acf[3 + n] = h->a_cfth[n].p;
      a.acth[n] = h->a_cth[n].p;
      a.pr[n] = p.pr[n];
      a.dthdx[n] = p.dthdx[n];
      a.dthdz[n] = p.dthdz[n];
      a.jst[n] = h->jstart_th[n];
      a.jet[n] = h->jend_th[n];
    }
    a.acu[0] = h->a_cu[0].p;
    a.acu[1] = h->a_cu[1].p;
    a.acu[2] = h->a_cu[2].p;
    a.acp = h->a_cp.p;
    a.kx = h->d_kx.p;
    a.kz = h->d_kz.p;
    a.dy = h->d_dy.p;
    a.dyf = h->d_dyf.p;
    a.nu = h->nu;
    a.beta = p.beta;
    a.coriolis = p.ro_inv / p.delta;
    a.grav_x = p.grav_x;
    a.grav_y = p.grav_y;
    a.grav_z = p.grav_z;
    a.temp3 = temp3;
    a.temp4 = temp4;
    a.temp5 = temp5;
    set_front(h, a);
    a.rk_step = rk_step;
    a.nth = nth;
    a.js = h->jstart;
    a.je = h->jend;
    Timed tm(h, DGPU_K_ADJOINT, 16.0 * (5.0 * nf + 8) * ns);
    hipLaunchKernelGGL(k_adj_rhs_a, grid1((long long)ns), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_adj_rhs_b, grid1((long long)ns), dim3(256), 0, h->stream, a, g);
    HIPCHECK(hipGetLastError());
  }
  // 6. product slots: Xbar = Fz^T Zbar, Zbar_s = -fbar or ikz fbar (the
  //    combinations of k_rhs: f1 -= Z0 + ikz Z1, f3 -= Z2 + ikz Z3,
  //    f2 -= Z4 + ikz Z5, ft -= Z6 + ikz Z7)
  //    (slot by slot, so that each source field's second read, for its
  //    ikz slot, follows its first and comes from the last-level cache:
  //    batched by operator, the even slots then the odd ones, the adjoint's
  //    z inverses took 5.14 -> 5.31 ms per step at 512x257x512)
  const int nslot = 6 + 2 * nth;
  for (int s = 0; s < nslot; ++s) {
    int src;
    if (s < 6)
      src = (s < 2) ? 0 : (s < 4 ? 2 : 1);
    else
      src = 3 + (s - 6) / 2;
    const bool dz = (s & 1) != 0;
    const double sc = (dz ? 1.0 : -1.0) / g.nz;
    launch_zinv_op<double>(h, h->a_fb[src].p, xspec<double>(h->tpool[s]), 0, n2 - 1, dz ? OP_DZ : OP_ID, sc, sc);
  }
  // 7. the forward fields at the substep's input (with LES: after its wall
  //    values, rebuilt by les_adj_recompute)
  if (!les) {  // each field through its own idle cr x-spectra buffer, all in one launch per pass
    const double2* C[3 + DGPU_MAX_TH];
    double* U[3 + DGPU_MAX_TH];
    cpx<double>* T[3 + DGPU_MAX_TH];
    for (int f = 0; f < nf; ++f) {
      C[f] = f < 3 ? h->cu[f].p : h->cth[f - 3].p;
      U[f] = f < 3 ? h->u[f].p : h->th[f - 3].p;
      T[f] = xspec<double>(h->tpool[nslot + f]);
    }
    to_physical_fields<double>(h, nf, C, U, T);
  }
  // 8. product adjoints (x pass) and their Jacobians
  {
    AdjPairArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int s = 0; s < nslot; ++s) a.X[s] = h->tpool[s].p;
    const int npairs = 4 + 2 * nth;
    for (int q = 0; q < npairs; ++q) {
      a.outA[q] = h->a_pA[q].p;
      a.outB[q] = h->a_pB[q].p;
    }
    a.kx = h->d_kx.p;
    a.npairs = npairs;
    // KT z lines per workgroup, one transform group per line
    constexpr int KT = XTile<double>::KT;
    const int ng = x_ngroups<double>(h, KT, 512);
    const size_t lds = x_lds<double>(h, ng);
    dim3 grid((g.nz + KT - 1) / KT, Ny), block(ng * (g.nx / 8));
    Timed tm(h, DGPU_K_ADJOINT, (16.0 * g.tp * nslot + 16.0 * g.pp * npairs) * Ny, 16.0 * g.tp * nslot * Ny);
    auto kern = g.nx <= 512 ? k_adj_xpairs<true> : k_adj_xpairs<false>;
    if (g.nx == 512 && use_xl<double>(h, XL_ADJ))
      kern = xl_powers<double>(h, XL_ADJ) ? k_adj_xpairs<true, false, FFT_XL512P> : k_adj_xpairs<true, false, FFT_XL512>;
    if (g.nz % KT) kern = g.nx <= 512 ? k_adj_xpairs<true, true> : k_adj_xpairs<false, true>;
    // the power-twiddle form keeps no table in LDS
    const bool twp = kern == k_adj_xpairs<true, false, FFT_XL512P>;
    hipLaunchKernelGGL(kern, grid, block, twp ? x_lds<double>(h, ng, false) : lds, h->stream, a, g, twx_of<double>(h), 1);
    HIPCHECK(hipGetLastError());
  }
  {
    AdjProdArgs a;
    std::memset(&a, 0, sizeof(a));
    a.u1 = h->u[0].p;
    a.u2 = h->u[1].p;
    a.u3 = h->u[2].p;
    for (int n = 0; n < nth; ++n) {
      a.th[n] = h->th[n].p;
      a.ubth[n] = h->a_ub[3 + n].p;
    }
    for (int q = 0; q < 4 + 2 * nth; ++q) {
      a.A[q] = h->a_pA[q].p;
      a.B[q] = h->a_pB[q].p;
    }
    a.ub1 = h->a_ub[0].p;
    a.ub2 = h->a_ub[1].p;
    a.ub3 = h->a_ub[2].p;
    a.dy = h->d_dy.p;
    a.dyf = h->d_dyf.p;
    a.nth = nth;
    a.store = spec ? 1 : 0;  // the spectral transposed solve wrote no ubar
    Timed tm(h, DGPU_K_ADJOINT, 8.0 * npx * (3 + nth + 2 * (4 + 2 * nth) + (spec ? 1 : 2) * (3 + nth)));
    if (nth <= 2) {  // chunks of 16 planes per thread, every value read once a chunk
      constexpr int J = 16;
      const dim3 grid((unsigned)((g.pp + 255) / 256), (unsigned)((n2 + J - 1) / J));
      auto kern = nth == 0 ? k_adj_prod_march<0> : nth == 1 ? k_adj_prod_march<1> : k_adj_prod_march<2>;
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, h->stream, a, g, J);
    } else {
      hipLaunchKernelGGL(k_adj_prod, grid1((long long)npx), dim3(256), 0, h->stream, a, g);
    }
    HIPCHECK(hipGetLastError());
  }
  if (les) les_adj_stage(h, rk_step);
  // 9. cubar += P^T ubar (the inverse transforms of the state)
  for (int f = 0; f < nf; ++f) {
    launch_xfwd<double>(h, h->a_ub[f].p, T0, 0, n2 - 1);
    launch_zfwd<double>(h, T0, les ? h->a_cub[f].p : adj_spec(h, f), 0, n2 - 1, true, NxNz, 2.0 * NxNz);
  }
  // 10. with LES, the transposed wall values (k_bc_post, k_les_bc), then the
  //     post-BC adjoint joins the pre-BC one
  if (les) {
    AdjBcArgs a;
    std::memset(&a, 0, sizeof(a));
    a.c1 = h->a_cub[0].p;
    a.c2 = h->a_cub[1].p;
    a.c3 = h->a_cub[2].p;
    a.nth = nth;
    for (int n = 0; n < nth; ++n) {
      a.cth[n] = h->a_cub[3 + n].p;
      a.th_lo[n] = p.th_bc_ymin[n];
      a.th_hi[n] = p.th_bc_ymax[n];
    }
    a.u_lo = p.u_bc_ymin;
    a.v_lo = p.v_bc_ymin;
    a.w_lo = p.w_bc_ymin;
    a.u_hi = p.u_bc_ymax;
    a.v_hi = p.v_bc_ymax;
    a.w_hi = p.w_bc_ymax;
    {
      Timed tm(h, DGPU_K_ADJOINT, 16.0 * 12 * g.sp * (3 + nth));
      hipLaunchKernelGGL(k_adj_les_bc, grid1(g.sp), dim3(256), 0, h->stream, a, g);
      HIPCHECK(hipGetLastError());
    }
    for (int f = 0; f < nf; ++f) {
      Timed tm(h, DGPU_K_ADJOINT, 48.0 * ns);
      hipLaunchKernelGGL(k_axpy2, dim3(2048), dim3(256), 0, h->stream, adj_spec(h, f), h->a_cub[f].p,
                         (long long)ns);
      HIPCHECK(hipGetLastError());
    }
  }
}

// Checkpoint slot `slot` of the forward state.  mode 1 saves the whole state,
// mode 2 only the fields the adjoint substep reads (cu1..3, cth), mode 0
// restores what the slot holds.
void checkpoint(dgpu_handle* h, int slot, int mode) {
  const size_t ns = (size_t)h->g.ny2 * h->g.sp;
  const int nth = h->nth;
  std::vector<double2*> live = {h->cu[0].p, h->cu[1].p, h->cu[2].p};
  for (int n = 0; n < nth; ++n) live.push_back(h->cth[n].p);
  const size_t nlight = live.size();
  live.push_back(h->cp.p);
  for (int f = 0; f < 3; ++f) live.push_back(h->cf[f].p);
  for (int n = 0; n < nth; ++n) live.push_back(h->cfth[n].p);
  if (slot < 0 || slot > 4096) throw ApiError{DGPU_EINVAL, "checkpoint slot out of range"};
  if (mode < 0 || mode > 2) throw ApiError{DGPU_EINVAL, "checkpoint mode must be 0, 1 or 2"};
  if ((size_t)slot >= h->snaps.size()) {
    if (mode == 0) throw ApiError{DGPU_EINVAL, "checkpoint slot was never saved"};
    h->snaps.resize(slot + 1);
  }
  auto& sn = h->snaps[slot];
  if (mode == 0 && sn.nsaved == 0) throw ApiError{DGPU_EINVAL, "checkpoint slot was never saved"};
  const size_t nsave = (mode == 2) ? nlight : live.size();
  if (mode != 0) {
    while (sn.f.size() < nsave) {
      sn.f.emplace_back();
      sn.f.back().alloc(ns);
    }
    sn.nsaved = nsave;
    for (size_t i = 0; i < nsave; ++i) copy_spec(h, sn.f[i].p, live[i]);
    sn.w_transient = h->w_transient;
  } else {
    for (size_t i = 0; i < sn.nsaved; ++i) copy_spec(h, live[i], sn.f[i].p);
    h->w_transient = sn.w_transient;
  }
}

// -------------------------------------------------- reference layouts ----
size_t ref_spec_doubles(const Geo& g) { return (size_t)(g.nx + 2) * (g.nz + 2) * (g.ny + 2); }

void spec_from_ref(const dgpu_handle* h, const double* ref, std::vector<double2>& out) {
  const Geo& g = h->g;
  out.assign((size_t)g.ny2 * g.sp, make_double2(0, 0));
  const size_t NXC = g.nx / 2 + 1, NZ2 = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int q = 0; q < g.nq; ++q)
      for (int kk = 0; kk < g.nkk; ++kk) {
        size_t c = q + NXC * (kk + NZ2 * j);
        out[(size_t)j * g.sp + (size_t)q * g.nkk + kk] = make_double2(ref[2 * c], ref[2 * c + 1]);
      }
}

void spec_to_ref(const dgpu_handle* h, const std::vector<double2>& in, double* ref) {
  const Geo& g = h->g;
  std::memset(ref, 0, sizeof(double) * ref_spec_doubles(g));
  const size_t NXC = g.nx / 2 + 1, NZ2 = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int q = 0; q < g.nq; ++q)
      for (int kk = 0; kk < g.nkk; ++kk) {
        size_t c = q + NXC * (kk + NZ2 * j);
        double2 v = in[(size_t)j * g.sp + (size_t)q * g.nkk + kk];
        ref[2 * c] = v.x;
        ref[2 * c + 1] = v.y;
      }
}

void phys_from_ref(const dgpu_handle* h, const double* ref, std::vector<double>& out) {
  const Geo& g = h->g;
  out.assign((size_t)g.ny2 * g.pp, 0.0);
  const size_t NX2 = g.nx + 2, NZ2 = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int k = 0; k < g.nz; ++k)
      for (int i = 0; i < g.nx; ++i)
        out[(size_t)j * g.pp + (size_t)k * g.nx + i] = ref[i + NX2 * (k + NZ2 * j)];
}

void phys_to_ref(const dgpu_handle* h, const std::vector<double>& in, double* ref) {
  const Geo& g = h->g;
  std::memset(ref, 0, sizeof(double) * ref_spec_doubles(g));
  const size_t NX2 = g.nx + 2, NZ2 = g.nz + 2;
  for (int j = 0; j < g.ny2; ++j)
    for (int k = 0; k < g.nz; ++k)
      for (int i = 0; i < g.nx; ++i)
        ref[i + NX2 * (k + NZ2 * j)] = in[(size_t)j * g.pp + (size_t)k * g.nx + i];
}

// ---------------------------------------- fields in caller device memory --
// The reference layouts on the device (dgpu_*_field_device): the same
// mappings as spec_from_ref / spec_to_ref / phys_from_ref / phys_to_ref, one
// thread per element.  The reference-layout side is written in full (zeros
// outside the dealiased block or the physical points), so the caller's
// buffer is fully defined.
template <typename S>
__global__ void __launch_bounds__(256) k_spec_from_ref(const double2* __restrict__ ref, S* __restrict__ dst,
                                                       Geo g) {
  const long long n = (long long)g.ny2 * g.sp;
  const long long nxc = g.nx / 2 + 1, nz2 = g.nz + 2;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int j = (int)(t / g.sp);
    const int m = (int)(t - (long long)j * g.sp);
    const int q = m / g.nkk, kk = m - q * g.nkk;
    const double2 v = ref[q + nxc * (kk + nz2 * j)];
    dst[t] = S{(decltype(S::x))v.x, (decltype(S::x))v.y};
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_spec_to_ref(const S* __restrict__ src, double2* __restrict__ ref,
                                                     Geo g) {
  const long long nxc = g.nx / 2 + 1, nz2 = g.nz + 2;
  const long long n = nxc * nz2 * g.ny2;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int j = (int)(t / (nxc * nz2));
    const long long r = t - (long long)j * nxc * nz2;
    const int kk = (int)(r / nxc), q = (int)(r - (long long)kk * nxc);
    double2 v = make_double2(0.0, 0.0);
    if (q < g.nq && kk < g.nkk) {
      const S z = src[(long long)j * g.sp + (long long)q * g.nkk + kk];
      v = make_double2((double)z.x, (double)z.y);
    }
    ref[t] = v;
  }
}

__global__ void __launch_bounds__(256) k_phys_from_ref(const double* __restrict__ ref, double* __restrict__ dst,
                                                       Geo g) {
  const long long n = (long long)g.ny2 * g.pp;
  const long long nx2 = g.nx + 2, nz2 = g.nz + 2;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int j = (int)(t / g.pp);
    const int m = (int)(t - (long long)j * g.pp);
    const int k = m / g.nx, i = m - k * g.nx;
    dst[t] = ref[i + nx2 * (k + nz2 * j)];
  }
}

__global__ void __launch_bounds__(256) k_phys_to_ref(const double* __restrict__ src, double* __restrict__ ref, Geo g) {
  const long long nx2 = g.nx + 2, nz2 = g.nz + 2;
  const long long n = nx2 * nz2 * g.ny2;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
    const int j = (int)(t / (nx2 * nz2));
    const long long r = t - (long long)j * nx2 * nz2;
    const int k = (int)(r / nx2), i = (int)(r - (long long)k * nx2);
    ref[t] = (i < g.nx && k < g.nz) ? src[(long long)j * g.pp + (long long)k * g.nx + i] : 0.0;
  }
}

bool field_buf(dgpu_handle* h, int f, int n, void** dev, bool* spec);
bool adj_field_buf(dgpu_handle* h, int f, int n, void** dev);
int nfield_copies(const dgpu_handle* h, int f);
bool carry32(const dgpu_handle* h, int f);
void adj_alloc(dgpu_handle* h);
size_t ref_spec_doubles(const Geo& g);

// Copy field `field` between the handle (forward or adjoint state) and the
// caller's device buffer `dev` in the reference layout, on the caller's
// stream `cs`.  The copy waits for the work already queued on the handle's
// stream, and the handle's later work waits for the copy; the host does not
// wait.
void field_device_copy(dgpu_handle* h, int field, void* dev, hipStream_t cs, bool to_handle, bool adjoint) {
  const Geo& g = h->g;
  if (adjoint) {
    if (field < 0 || field > DGPU_CFTH) throw ApiError{DGPU_EINVAL, "bad adjoint field id"};
    adj_alloc(h);
  } else if (field < 0 || field >= DGPU_NFIELDS) {
    throw ApiError{DGPU_EINVAL, "bad field id"};
  }
  if (!h->ev_in) {
    HIPCHECK(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
  }
  HIPCHECK(hipEventRecord(h->ev_in, h->stream));
  HIPCHECK(hipStreamWaitEvent(cs, h->ev_in, 0));
  const size_t per = ref_spec_doubles(g);
  const dim3 grid(2048), block(256);
  for (int n = 0; n < nfield_copies(h, field); ++n) {
    void* buf = nullptr;
    bool spec = true;
    if (adjoint)
      adj_field_buf(h, field, n, &buf);
    else
      field_buf(h, field, n, &buf, &spec);
    double* ref = static_cast<double*>(dev) + n * per;
    const bool c32 = !adjoint && carry32(h, field);
    if (spec && to_handle && c32)
      hipLaunchKernelGGL(k_spec_from_ref<float2>, grid, block, 0, cs, reinterpret_cast<const double2*>(ref),
                         static_cast<float2*>(buf), g);
    else if (spec && to_handle)
      hipLaunchKernelGGL(k_spec_from_ref<double2>, grid, block, 0, cs, reinterpret_cast<const double2*>(ref),
                         static_cast<double2*>(buf), g);
    else if (spec && c32)
      hipLaunchKernelGGL(k_spec_to_ref<float2>, grid, block, 0, cs, static_cast<const float2*>(buf),
                         reinterpret_cast<double2*>(ref), g);
    else if (spec)
      hipLaunchKernelGGL(k_spec_to_ref<double2>, grid, block, 0, cs, static_cast<const double2*>(buf),
                         reinterpret_cast<double2*>(ref), g);
    else if (to_handle)
      hipLaunchKernelGGL(k_phys_from_ref, grid, block, 0, cs, ref, static_cast<double*>(buf), g);
    else
      hipLaunchKernelGGL(k_phys_to_ref, grid, block, 0, cs, static_cast<const double*>(buf), ref, g);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(h->ev_out, cs));
  HIPCHECK(hipStreamWaitEvent(h->stream, h->ev_out, 0));
}

// (buffer, is_spectral) for field id f, scalar n
bool adj_field_buf(dgpu_handle* h, int f, int n, void** dev) {
  switch (f) {
    case DGPU_CU1: *dev = h->a_cu[0].p; return true;
    case DGPU_CU2: *dev = h->a_cu[1].p; return true;
    case DGPU_CU3: *dev = h->a_cu[2].p; return true;
    case DGPU_CP: *dev = h->a_cp.p; return true;
    case DGPU_CTH: *dev = h->a_cth[n].p; return true;
    case DGPU_CF1: *dev = h->a_cf[0].p; return true;
    case DGPU_CF2: *dev = h->a_cf[1].p; return true;
    case DGPU_CF3: *dev = h->a_cf[2].p; return true;
    case DGPU_CFTH: *dev = h->a_cfth[n].p; return true;
  }
  return false;
}

bool field_buf(dgpu_handle* h, int f, int n, void** dev, bool* spec) {
  *spec = true;
  switch (f) {
    case DGPU_CU1: *dev = h->cu[0].p; return true;
    case DGPU_CU2: *dev = h->cu[1].p; return true;
    case DGPU_CU3: *dev = h->cu[2].p; return true;
    case DGPU_CP: *dev = h->cp.p; return true;
    case DGPU_CTH: *dev = h->cth[n].p; return true;
    case DGPU_CF1: *dev = h->cf[0].p; return true;
    case DGPU_CF2: *dev = h->cf[1].p; return true;
    case DGPU_CF3: *dev = h->cf[2].p; return true;
    case DGPU_CFTH: *dev = h->cfth[n].p; return true;
    case DGPU_NU_T: *dev = h->nu_t.p; *spec = false; return true;
    case DGPU_KAPPA_T: *dev = h->kappa_t[n].p; *spec = false; return true;
  }
  return false;
}

int nfield_copies(const dgpu_handle* h, int f) {
  return (f == DGPU_CTH || f == DGPU_CFTH || f == DGPU_KAPPA_T) ? h->nth : 1;
}

// The RK carry cf1..3, cfth is stored in the transform precision: FP32
// spectra in the mixed form (k_rhs), FP64 otherwise.
bool carry32(const dgpu_handle* h, int f) {
  return h->precision == DGPU_MIXED && f >= DGPU_CF1 && f <= DGPU_CFTH;
}

int check_params(const dgpu_params* p, std::string& why) {
  auto pow2 = [](int n) { return n >= 16 && n <= 4096 && (n & (n - 1)) == 0; };
  if (!pow2(p->nx)) { why = "nx must be a power of two in [16, 4096]"; return -1; }
  if (p->nz < 16 || p->nz > 2048) { why = "nz must be in [16, 2048]"; return -1; }
  if (p->ny < 5) { why = "ny must be >= 5"; return -1; }
  if (p->n_th < 0 || p->n_th > DGPU_MAX_TH) { why = "n_th out of range"; return -1; }
  if (!(p->re > 0) || !(p->delta_t > 0)) { why = "re and delta_t must be positive"; return -1; }
  if (p->delta == 0.0 && p->ro_inv != 0.0) { why = "delta must be nonzero when ro_inv is"; return -1; }
  for (int n = 0; n < p->n_th; ++n)
    if (!(p->pr[n] > 0)) { why = "pr must be positive"; return -1; }
  if (p->use_les && p->les_model_type != 1 && p->les_model_type != 4 && p->les_model_type != 5) {
    why = "les_model_type must be 1 (Smagorinsky), 4 (AMD-Rozema) or 5 (AMD-Verstappen)";
    return -1;
  }
  auto bc01 = [](int b) { return b == 0 || b == 1; };
  if (!bc01(p->u_bc_ymin) || !bc01(p->u_bc_ymax) || !bc01(p->v_bc_ymin) || !bc01(p->v_bc_ymax) ||
      !bc01(p->w_bc_ymin) || !bc01(p->w_bc_ymax)) {
    why = "velocity BC types must be 0 (Dirichlet) or 1 (Neumann)";
    return -1;
  }
  for (int n = 0; n < p->n_th; ++n)
    if (!bc01(p->th_bc_ymin[n]) || !bc01(p->th_bc_ymax[n])) {
      why = "scalar BC types must be 0 or 1";
      return -1;
    }
  return 0;
}

void free_all(dgpu_handle* h) {
  for (int f = 0; f < 3; ++f) {
    h->cu[f].release();
    h->cf[f].release();
    h->u[f].release();
    h->r[f].release();
  }
  h->cp.release();
  for (int n = 0; n < DGPU_MAX_TH; ++n) {
    h->cth[n].release();
    h->cfth[n].release();
    h->th[n].release();
    h->rth[n].release();
    h->kappa_t[n].release();
  }
  h->nu_t.release();
  for (auto& b : h->tpool) b.release();
  h->tpool.clear();
  for (auto& b : h->zpool) b.release();
  h->zpool.clear();
  h->sb.release();
  h->sg.release();
  h->rd_ib.release();
  h->d_prow.release();
  h->twx.release();
  h->twz.release();
  h->twx32.release();
  h->twz32.release();
  h->d_kx.release();
  h->d_kz.release();
  h->d_dy.release();
  h->d_dyf.release();
  h->d_gyf.release();
  h->d_front.release();
  h->d_flags.release();
  h->pois.release();
  h->cfl_cdy.release();
  h->cfl_offx.release();
  h->cfl_offz.release();
  h->cfl_part.release();
  h->rows.release();
  h->trirows.release();
  for (auto& b : h->grad) b.release();
  for (auto& b : h->tau) b.release();
  for (auto& b : h->ctau) b.release();
  for (auto& b : h->lesf) b.release();
  for (auto& b : h->wpre) b.release();
  for (int n = 0; n < DGPU_MAX_TH; ++n) {
    for (int w = 0; w < 3; ++w) h->dth[w][n].release();
    for (int w = 0; w < 2; ++w) {
      h->thf[w][n].release();
      h->cthf[w][n].release();
    }
    h->lesfth[n].release();
  }
  h->sbl.release();
  h->sgl.release();
  for (auto& b : h->a_cu) b.release();
  for (auto& b : h->a_cf) b.release();
  for (auto& b : h->a_cth) b.release();
  for (auto& b : h->a_cfth) b.release();
  for (auto& b : h->a_cr) b.release();
  for (auto& b : h->a_fb) b.release();
  for (auto& b : h->a_rb) b.release();
  for (auto& b : h->a_ub) b.release();
  for (auto& b : h->a_pA) b.release();
  for (auto& b : h->a_pB) b.release();
  h->a_cp.release();
  h->rowsT.release();
  h->d_w.release();
  for (auto& b : h->cap_cub) b.release();
  for (auto& b : h->cap_x) b.release();
  for (auto& b : h->fsave) b.release();
  h->fsave.clear();
  for (auto& b : h->a_cub) b.release();
  h->a_nt.release();
  for (auto& b : h->a_kt) b.release();
  for (auto& b : h->a_G) b.release();
  for (int w = 0; w < 3; ++w)
    for (auto& b : h->a_T[w]) b.release();
  for (auto& sn : h->snaps)
    for (auto& b : sn.f) b.release();
  h->snaps.clear();
}

void create_impl(dgpu_handle* h, const double* gy) {
  const dgpu_params& p = h->p;
  Geo& g = h->g;
  g.nx = p.nx;
  g.ny = p.ny;
  g.nz = p.nz;
  g.lognx = 0;
  while ((1 << g.lognx) < g.nx) ++g.lognx;
  g.lognz = 0;
  while ((1 << g.lognz) < g.nz) ++g.lognz;
  h->zgen = (1 << g.lognz) != g.nz;
  if (h->zgen) h->zplan = make_zplan(g.nz);
  g.nkx = p.nx / 3;
  g.nkz = p.nz / 3;
  g.nq = g.nkx + 1;
  g.nkk = 2 * g.nkz + 1;
  g.ny2 = p.ny + 2;
  g.pp = (long long)g.nx * g.nz;
  g.sp = (long long)g.nq * g.nkk;
  g.tp = (long long)g.nq * g.nz;
  h->nth = p.n_th;
  h->nu = 1.0 / p.re;  // parameters.f90:101
  setup_grid(h, gy);

  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, dev));
  h->lds_limit = prop.sharedMemPerBlock;
  if (h->lds_limit > 65536) {
    const void* kernels[] = {
        (const void*)k_rhs<double, true>,     (const void*)k_rhs<double, false>,
        (const void*)k_xfwd<double, true>,    (const void*)k_xfwd<double, false>,
        (const void*)k_xinv<double, true>,    (const void*)k_xinv<double, false>,
        (const void*)k_zfwd<double, true>,    (const void*)k_zfwd<double, false>,
        (const void*)k_zinv<double, true>,    (const void*)k_zinv<double, false>,
        (const void*)k_zinv_op<double, true>, (const void*)k_zinv_op<double, false>,
        (const void*)k_xprod_all<double, true>, (const void*)k_xprod_all<double, false>,
        (const void*)k_xprod_all<double, true, 9>, (const void*)k_rhs<double, true, 9>,
        (const void*)k_xinv<double, true, 9>,     (const void*)k_zinv<double, true, 9>,
        (const void*)k_rhs<float, true, 9>,
        (const void*)k_rhs<float, true>,      (const void*)k_rhs<float, false>,
        (const void*)k_xfwd<float, true>,     (const void*)k_xfwd<float, false>,
        (const void*)k_xinv<float, true>,     (const void*)k_xinv<float, false>,
        (const void*)k_zfwd<float, true>,     (const void*)k_zfwd<float, false>,
        (const void*)k_zinv<float, true>,     (const void*)k_zinv<float, false>,
        (const void*)k_zinv_op<float, true>,  (const void*)k_zinv_op<float, false>,
        (const void*)k_xprod_all<float, true>, (const void*)k_xprod_all<float, false>,
        (const void*)k_adj_xpairs<true>,      (const void*)k_adj_xpairs<false>,
        (const void*)k_xfwd<double, false, true>, (const void*)k_xinv<double, false, 0, true>,
        (const void*)k_xprod_all<double, false, 0, true>, (const void*)k_adj_xpairs<false, true>,
        (const void*)k_xfwd<float, false, true>,  (const void*)k_xinv<float, false, 0, true>,
        (const void*)k_xprod_all<float, false, 0, true>,
        (const void*)k_xfwd<double, true, true>,  (const void*)k_xinv<double, true, 0, true>,
        (const void*)k_xprod_all<double, true, 0, true>, (const void*)k_adj_xpairs<true, true>,
        (const void*)k_xfwd<float, true, true>,   (const void*)k_xinv<float, true, 0, true>,
        (const void*)k_xprod_all<float, true, 0, true>,
        (const void*)k_xprod_all<double, true, 0, false, true>, (const void*)k_xprod_all<double, false, 0, false, true>,
        (const void*)k_xprod_all<double, true, 9, false, true>,
        (const void*)k_xprod_all<double, true, FFT_XL512, false, true>, (const void*)k_xprod_all<float, true, FFT_XL512>,
        (const void*)k_rhs<double, true, FFT_XL512>,   (const void*)k_rhs<float, true, FFT_XL512>,
        (const void*)k_xinv<double, true, FFT_XL512>,  (const void*)k_xinv<float, true, FFT_XL512>,
        (const void*)k_zinv<double, true, FFT_XL512>,  (const void*)k_zinv<float, true, FFT_XL512>,
        (const void*)k_zinv<double, true, FFT_XL512P>, (const void*)k_zfwd<double, true, FFT_XL512P>,
        (const void*)k_zinv_op<double, true, FFT_XL512P>,
        (const void*)k_xfwd<double, true, false, FFT_XL512>, (const void*)k_xfwd<float, true, false, FFT_XL512>,
        (const void*)k_zfwd<double, true, FFT_XL512>,  (const void*)k_zfwd<float, true, FFT_XL512>,
        (const void*)k_zinv_op<double, true, FFT_XL512>, (const void*)k_zinv_op<float, true, FFT_XL512>,
        (const void*)k_adj_xpairs<true, false, FFT_XL512>, (const void*)k_adj_xpairs<true, false, FFT_XL512P>,
        (const void*)k_xprod_all<double, true, FFT_XL512P, false, true>, (const void*)k_xinv<double, true, FFT_XL512P>,
        (const void*)k_xprod_all<float, true, FFT_XL512, false, true>,
        (const void*)k_xfwd<double, true, false, FFT_XL512P>,
        (const void*)k_xprod_all<double, true, 0, true, true>, (const void*)k_xprod_all<double, false, 0, true, true>,
        (const void*)k_zgen_fwd<double>, (const void*)k_zgen_inv<double>,
        (const void*)k_zgen_fwd<float>,  (const void*)k_zgen_inv<float>};
    for (const void* k : kernels)
      HIPCHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds_limit));
  }
  HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->scratch = scratch_address();

  // kernels index fields with 32-bit arithmetic (plane_of)
  if ((double)g.ny2 * (double)g.pp >= 4294967296.0)
    throw ApiError{DGPU_EUNSUP, "grid too large: a field must have fewer than 2^32 points"};
  const size_t ns = (size_t)g.ny2 * g.sp, np = (size_t)g.ny2 * g.pp, nt = (size_t)g.ny2 * g.tp;
  for (int f = 0; f < 3; ++f) {
    h->cu[f].alloc(ns);
    h->cf[f].alloc(ns);
    h->u[f].alloc(np);
    h->r[f].alloc(np);
  }
  h->cp.alloc(ns);
  for (int n = 0; n < h->nth; ++n) {
    h->cth[n].alloc(ns);
    h->cfth[n].alloc(ns);
    h->th[n].alloc(np);
    h->rth[n].alloc(np);
    h->kappa_t[n].alloc(np);
  }
  h->nu_t.alloc(np);
  const int nslot = 6 + 2 * h->nth, ncr = 3 + h->nth;
  h->tpool.resize(nslot + ncr);
  for (auto& b : h->tpool) b.alloc(nt);
  if (h->zgen) {
    h->zpool.resize(nslot);
    for (auto& b : h->zpool) b.alloc(ns);
  }
  h->sb.alloc(ns);
  h->sg.alloc(ns);
  for (int f = 0; f < 3; ++f) {
    HIPCHECK(hipMemset(h->cu[f].p, 0, sizeof(double2) * ns));
    HIPCHECK(hipMemset(h->cf[f].p, 0, sizeof(double2) * ns));
    HIPCHECK(hipMemset(h->u[f].p, 0, sizeof(double) * np));
  }
  HIPCHECK(hipMemset(h->cp.p, 0, sizeof(double2) * ns));
  for (int n = 0; n < h->nth; ++n) {
    HIPCHECK(hipMemset(h->cth[n].p, 0, sizeof(double2) * ns));
    HIPCHECK(hipMemset(h->cfth[n].p, 0, sizeof(double2) * ns));
    HIPCHECK(hipMemset(h->kappa_t[n].p, 0, sizeof(double) * np));
  }
  HIPCHECK(hipMemset(h->nu_t.p, 0, sizeof(double) * np));

  std::vector<double2> w;
  std::vector<float2> w32;
  auto to32 = [&]() {
    w32.resize(w.size());
    for (size_t t = 0; t < w.size(); ++t) w32[t] = make_float2((float)w[t].x, (float)w[t].y);
  };
  twiddles(g.nx, w);
  upload(h->twx, w);
  to32();
  upload(h->twx32, w32);
  twiddles(g.nz, w);
  upload(h->twz, w);
  to32();
  upload(h->twz32, w32);
  upload(h->d_kx, h->kx);
  upload(h->d_kz, h->kz);
  upload(h->d_dy, h->dy);
  upload(h->d_dyf, h->dyf);
  upload(h->d_gyf, h->gyf);
  if (p.flavor_front && p.ro_inv != 0.0 && h->nth > 0) {
    const int n2 = g.ny2;
    std::vector<double> fr(2 * (size_t)n2, 0.0);
    const double dr = p.delta / p.ro_inv;
    for (int j = 0; j < n2; ++j) fr[j] = (h->gyf[j] - 0.5 * p.ly) * dr;
    for (int j = 1; j < n2; ++j) fr[n2 + j] = (h->gy[j] - 0.5 * p.ly) * dr;
    upload(h->d_front, fr);
  }
  setup_projection(h);

  // the x passes and the products pass are compiled for 512-lane workgroups
  for (int t : {h->kn.xpass_threads, h->kn.xprod_threads, h->kn.xprod_threads32})
    if (t < 64 || t > 512)
      throw ApiError{DGPU_EINVAL, "DGPU_XPASS_THREADS and DGPU_XPROD_THREADS(32) must lie in [64, 512]"};
  // CN factor tables (DGPU_SEG_TARGET overrides the target segment length)
  if (h->kn.seg_target < 1) throw ApiError{DGPU_EINVAL, "DGPU_SEG_TARGET must be a positive row count"};
  const int n = g.ny2;
  h->seg_L = seg_len_for(n, h->kn.seg_target);
  if (h->seg_L < 0) throw ApiError{DGPU_EUNSUP, "ny too large for the segmented solver"};
  h->seg_S = (n + h->seg_L - 1) / h->seg_L;
  h->nfields = 3 + h->nth;
  build_cn_tables(h);
  if (p.use_les) {
    for (int m = 0; m < 9; ++m) h->grad[m].alloc(np);
    for (int m = 0; m < 7; ++m) {
      h->tau[m].alloc(np);
      h->ctau[m].alloc(ns);
    }
    for (int f = 0; f < 3; ++f) {
      h->lesf[f].alloc(ns);
      h->wpre[f].alloc((size_t)4 * g.sp);
    }
    for (int nn = 0; nn < h->nth; ++nn) {
      for (int w = 0; w < 3; ++w) h->dth[w][nn].alloc(np);
      for (int w = 0; w < 2; ++w) {
        h->thf[w][nn].alloc(np);
        h->cthf[w][nn].alloc(ns);
      }
      h->lesfth[nn].alloc(ns);
      h->wpre[3 + nn].alloc((size_t)4 * g.sp);
    }
    // b', g' of the LES CN solves, one region per field (the fields of a
    // substep are solved in one launch)
    h->sbl.alloc(np * (3 + h->nth));
    h->sgl.alloc(np * (3 + h->nth));
  }
  HIPCHECK(hipDeviceSynchronize());
}

}  // namespace

// ================================================================ C-ABI ====

namespace {
int run_guarded(dgpu_handle* h, const std::function<void()>& fn) {
  try {
    fn();
    return DGPU_OK;
  } catch (const HipError& e) {
    if (h) h->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return e.e == hipErrorOutOfMemory ? DGPU_ENOMEM : DGPU_EHIP;
  } catch (const ApiError& e) {
    if (h) h->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (h) h->err = "host allocation failed";
    return DGPU_ENOMEM;
  } catch (...) {
    if (h) h->err = "unknown error";
    return DGPU_EINVAL;
  }
}
}  // namespace

extern "C" {

int dgpu_create(const dgpu_params* p, const double* gy, int precision, dgpu_handle** out) {
  if (!p || !gy || !out) return DGPU_EINVAL;
  *out = nullptr;
  std::string why;
  if (check_params(p, why)) return DGPU_EINVAL;
  if (precision != DGPU_FP64 && precision != DGPU_MIXED) return DGPU_EINVAL;
  if (precision == DGPU_MIXED && p->nz < XTile<float>::KT) return DGPU_EUNSUP;
  dgpu_handle* h = new (std::nothrow) dgpu_handle();
  if (!h) return DGPU_ENOMEM;
  h->p = *p;
  h->precision = precision;
  h->kn = knobs_from_env();
  int rc = run_guarded(h, [&] { create_impl(h, gy); });
  if (rc != DGPU_OK) {
    free_all(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return rc;
  }
  *out = h;
  return DGPU_OK;
}

int dgpu_destroy(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  (void)hipStreamSynchronize(h->stream);
  for (auto& m : h->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (auto e : h->event_pool) (void)hipEventDestroy(e);
  if (h->ev_in) (void)hipEventDestroy(h->ev_in);
  if (h->ev_out) (void)hipEventDestroy(h->ev_out);
  free_all(h);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return DGPU_OK;
}

const char* dgpu_last_error(const dgpu_handle* h) { return h ? h->err.c_str() : "null handle"; }

int64_t dgpu_field_doubles(const dgpu_handle* h, int field) {
  if (!h || field < 0 || field >= DGPU_NFIELDS) return -1;
  return (int64_t)ref_spec_doubles(h->g) * nfield_copies(h, field);
}

int dgpu_set_field(dgpu_handle* h, int field, const double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (field < 0 || field >= DGPU_NFIELDS) {
    h->err = "bad field id";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      bool spec;
      field_buf(h, field, n, &dev, &spec);
      if (spec && carry32(h, field)) {
        std::vector<double2> v;
        spec_from_ref(h, host + n * per, v);
        std::vector<float2> w(v.size());
        for (size_t t = 0; t < v.size(); ++t) w[t] = make_float2((float)v[t].x, (float)v[t].y);
        HIPCHECK(hipMemcpyAsync(dev, w.data(), sizeof(float2) * w.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
      } else if (spec) {
        std::vector<double2> v;
        spec_from_ref(h, host + n * per, v);
        HIPCHECK(hipMemcpyAsync(dev, v.data(), sizeof(double2) * v.size(), hipMemcpyHostToDevice,
                                h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
      } else {
        std::vector<double> v;
        phys_from_ref(h, host + n * per, v);
        HIPCHECK(hipMemcpyAsync(dev, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice,
                                h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
      }
    }
  });
}

int dgpu_get_field(dgpu_handle* h, int field, double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (field < 0 || field >= DGPU_NFIELDS) {
    h->err = "bad field id";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      bool spec;
      field_buf(h, field, n, &dev, &spec);
      if (spec && carry32(h, field)) {
        std::vector<float2> w((size_t)h->g.ny2 * h->g.sp);
        HIPCHECK(hipMemcpyAsync(w.data(), dev, sizeof(float2) * w.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        std::vector<double2> v(w.size());
        for (size_t t = 0; t < w.size(); ++t) v[t] = make_double2(w[t].x, w[t].y);
        spec_to_ref(h, v, host + n * per);
      } else if (spec) {
        std::vector<double2> v((size_t)h->g.ny2 * h->g.sp);
        HIPCHECK(hipMemcpyAsync(v.data(), dev, sizeof(double2) * v.size(), hipMemcpyDeviceToHost,
                                h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        spec_to_ref(h, v, host + n * per);
      } else {
        std::vector<double> v((size_t)h->g.ny2 * h->g.pp);
        HIPCHECK(hipMemcpyAsync(v.data(), dev, sizeof(double) * v.size(), hipMemcpyDeviceToHost,
                                h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        phys_to_ref(h, v, host + n * per);
      }
    }
  });
}

// Physical-space copies of the spectral state fields (fft_xz_to_physical /
// fft_xz_to_fourier around the flow-file I/O, phdf5.f90:1076-1082, 1190-1195,
// 1437-1446), computed with the FP64 transforms through the handle's scratch.
static bool physical_field_ok(dgpu_handle* h, int field) {
  if (field == DGPU_CU1 || field == DGPU_CU2 || field == DGPU_CU3 || field == DGPU_CP) return true;
  if (field == DGPU_CTH && h->nth > 0) return true;
  h->err = "physical copies exist for cu1, cu2, cu3, cp and cth only";
  return false;
}

int dgpu_get_physical(dgpu_handle* h, int field, double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (!physical_field_ok(h, field)) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      bool spec;
      field_buf(h, field, n, &dev, &spec);
      to_physical<double>(h, (const double2*)dev, h->r[0].p, xspec<double>(h->tpool[0]));
      std::vector<double> v((size_t)h->g.ny2 * h->g.pp);
      HIPCHECK(hipMemcpyAsync(v.data(), h->r[0].p, sizeof(double) * v.size(), hipMemcpyDeviceToHost,
                              h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
      phys_to_ref(h, v, host + n * per);
    }
  });
}

int dgpu_set_physical(dgpu_handle* h, int field, const double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (!physical_field_ok(h, field)) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      bool spec;
      field_buf(h, field, n, &dev, &spec);
      std::vector<double> v;
      phys_from_ref(h, host + n * per, v);
      HIPCHECK(hipMemcpyAsync(h->r[0].p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice,
                              h->stream));
      to_fourier<double>(h, h->r[0].p, (double2*)dev, xspec<double>(h->tpool[0]), 0, h->g.ny2 - 1);
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
  });
}

int dgpu_poisson_p(dgpu_handle* h, const double* ri) {
  if (!h || (h->nth > 0 && !ri)) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    poisson_p(h, ri);
    HIPCHECK(hipStreamSynchronize(h->stream));
  });
}

int dgpu_set_courant(dgpu_handle* h, const dgpu_courant_params* c) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { set_courant_impl(h, c); });
}

int dgpu_courant(dgpu_handle* h, double time) {
  if (!h) return DGPU_EINVAL;
  if (!h->var_dt) {
    h->err = "dgpu_courant needs the settings of dgpu_set_courant";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    const double pts = cfl_min(h, h->cu[0].p, h->cu[1].p, h->cu[2].p, true);
    set_delta_t_impl(h, courant_dt(h, pts, time));
  });
}

int dgpu_get_delta_t(dgpu_handle* h, double* delta_t) {
  if (!h || !delta_t) return DGPU_EINVAL;
  *delta_t = h->p.delta_t;
  return DGPU_OK;
}

int dgpu_set_delta_t(dgpu_handle* h, double delta_t) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { set_delta_t_impl(h, delta_t); });
}

int dgpu_apply_bc_vel_post(dgpu_handle* h, double time) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { launch_bc_post(h, time); });
}

int dgpu_rem_div(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { launch_remdiv(h, false, 1.0); });
}

int dgpu_rk_substep(dgpu_handle* h, int rk_step, double time, int time_step) {
  if (!h) return DGPU_EINVAL;
  if (rk_step < 1 || rk_step > 3) {
    h->err = "rk_step must be 1, 2 or 3";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] { rk_substep(h, rk_step, time, time_step); });
}

int dgpu_step(dgpu_handle* h, int nsteps, double* time, int32_t* time_step) {
  if (!h || !time || !time_step || nsteps < 0) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    for (int s = 0; s < nsteps; ++s) {
      *time_step += 1;
      for (int rk = 1; rk <= 3; ++rk) rk_substep(h, rk, *time, *time_step);
      *time += h->p.delta_t;  // diablo.f90:108
    }
  });
}

int dgpu_timing_enable(dgpu_handle* h, int enable) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    collect_timing(h);
    h->timing = enable != 0;
    for (int c = 0; c < DGPU_NCLASS; ++c) {
      h->t_ms[c] = 0.0;
      h->t_bytes[c] = 0.0;
      h->t_tbytes[c] = 0.0;
      h->t_count[c] = 0;
    }
  });
}

int dgpu_timing_read(dgpu_handle* h, double* ms, double* bytes, int64_t* count) {
  if (!h || !ms || !bytes || !count) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    collect_timing(h);
    for (int c = 0; c < DGPU_NCLASS; ++c) {
      ms[c] = h->t_ms[c];
      bytes[c] = h->t_bytes[c];
      count[c] = h->t_count[c];
    }
  });
}

int dgpu_timing_read_intermediate(dgpu_handle* h, double* tbytes) {
  if (!h || !tbytes) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    collect_timing(h);
    for (int c = 0; c < DGPU_NCLASS; ++c) tbytes[c] = h->t_tbytes[c];
  });
}

const char* dgpu_kernel_class_name(int cls) {
  static const char* names[DGPU_NCLASS] = {"x_forward", "x_inverse", "z_forward", "z_inverse",
                                           "x_products", "spectral_rhs", "cn_solve",
                                           "projection", "wall_bc", "les_model", "adjoint",
                                           "state_copy", "init"};
  return (cls >= 0 && cls < DGPU_NCLASS) ? names[cls] : "";
}

int dgpu_adj_set_field(dgpu_handle* h, int field, const double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (field < 0 || field > DGPU_CFTH) {
    h->err = "bad adjoint field id";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    adj_alloc(h);
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      adj_field_buf(h, field, n, &dev);
      std::vector<double2> v;
      spec_from_ref(h, host + n * per, v);
      HIPCHECK(hipMemcpyAsync(dev, v.data(), sizeof(double2) * v.size(), hipMemcpyHostToDevice, h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
  });
}

int dgpu_adj_get_field(dgpu_handle* h, int field, double* host) {
  if (!h || !host) return DGPU_EINVAL;
  if (field < 0 || field > DGPU_CFTH) {
    h->err = "bad adjoint field id";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    adj_alloc(h);
    const size_t per = ref_spec_doubles(h->g);
    for (int n = 0; n < nfield_copies(h, field); ++n) {
      void* dev;
      adj_field_buf(h, field, n, &dev);
      std::vector<double2> v((size_t)h->g.ny2 * h->g.sp);
      HIPCHECK(hipMemcpyAsync(v.data(), dev, sizeof(double2) * v.size(), hipMemcpyDeviceToHost, h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
      spec_to_ref(h, v, host + n * per);
    }
  });
}

int dgpu_set_field_device(dgpu_handle* h, int field, const double* dev, void* stream) {
  if (!h || !dev) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    field_device_copy(h, field, const_cast<double*>(dev), static_cast<hipStream_t>(stream), true, false);
  });
}

int dgpu_get_field_device(dgpu_handle* h, int field, double* dev, void* stream) {
  if (!h || !dev) return DGPU_EINVAL;
  return run_guarded(h, [&] { field_device_copy(h, field, dev, static_cast<hipStream_t>(stream), false, false); });
}

int dgpu_adj_set_field_device(dgpu_handle* h, int field, const double* dev, void* stream) {
  if (!h || !dev) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    field_device_copy(h, field, const_cast<double*>(dev), static_cast<hipStream_t>(stream), true, true);
  });
}

int dgpu_adj_get_field_device(dgpu_handle* h, int field, double* dev, void* stream) {
  if (!h || !dev) return DGPU_EINVAL;
  return run_guarded(h, [&] { field_device_copy(h, field, dev, static_cast<hipStream_t>(stream), false, true); });
}

int dgpu_adj_zero(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] {
    adj_alloc(h);
    const size_t ns = (size_t)h->g.ny2 * h->g.sp;
    for (int f = DGPU_CU1; f <= DGPU_CFTH; ++f)
      for (int n = 0; n < nfield_copies(h, f); ++n) {
        void* dev;
        adj_field_buf(h, f, n, &dev);
        HIPCHECK(hipMemsetAsync(dev, 0, sizeof(double2) * ns, h->stream));
      }
  });
}

int dgpu_adj_rk_substep(dgpu_handle* h, int rk_step, double time, int time_step) {
  if (!h) return DGPU_EINVAL;
  if (rk_step < 1 || rk_step > 3) {
    h->err = "rk_step must be 1, 2 or 3";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] { adj_substep(h, rk_step, time, time_step); });
}

int dgpu_adj_add_mean(dgpu_handle* h, int field, const double* w) {
  if (!h || !w) return DGPU_EINVAL;
  if (field != DGPU_CU1 && field != DGPU_CU2 && field != DGPU_CU3 && field != DGPU_CP) {
    h->err = "adj_add_mean takes cu1, cu2, cu3 or cp";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    adj_alloc(h);
    void* dev;
    adj_field_buf(h, field, 0, &dev);
    HIPCHECK(hipMemcpyAsync(h->d_w.p, w, sizeof(double) * h->g.ny2, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_add_mean, dim3((h->g.ny2 + 63) / 64), dim3(64), 0, h->stream,
                       reinterpret_cast<double2*>(dev), h->d_w.p, h->g);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(h->stream));
  });
}

int dgpu_mean_profile(dgpu_handle* h, int field, double* out) {
  if (!h || !out) return DGPU_EINVAL;
  if (field != DGPU_CU1 && field != DGPU_CU2 && field != DGPU_CU3 && field != DGPU_CP) {
    h->err = "mean_profile takes cu1, cu2, cu3 or cp";
    return DGPU_EINVAL;
  }
  return run_guarded(h, [&] {
    void* dev;
    bool spec;
    field_buf(h, field, 0, &dev, &spec);
    // the (0,0) entry of every plane: a strided copy of one double per plane
    HIPCHECK(hipMemcpy2DAsync(out, sizeof(double), dev, sizeof(double2) * h->g.sp, sizeof(double), h->g.ny2,
                              hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
  });
}

int dgpu_checkpoint(dgpu_handle* h, int slot, int save) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { checkpoint(h, slot, save); });
}

int dgpu_check_state(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { check_state_impl(h, false); });
}

int dgpu_check_state_noalloc(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { check_state_impl(h, true); });
}

int dgpu_synchronize(dgpu_handle* h) {
  if (!h) return DGPU_EINVAL;
  return run_guarded(h, [&] { HIPCHECK(hipStreamSynchronize(h->stream)); });
}

}  // extern "C"
