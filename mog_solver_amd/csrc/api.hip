// C ABI of the MOG force step (declared in include/mog.h).
//
// Owns the per-context device state: packed transforms, int64 density grids,
// complex scratch, float4 force grids and the FP64-derived tables, and
// sequences the kernels of one step on the caller's stream.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <complex>
#include <string>
#include <vector>

#include "../../include/mog.h"
#include "mog_common.h"

using namespace mog;

struct mog_ctx {
  unsigned long long uid;  // distinct for every context of the process
  mog_config cfg;
  Params P;
  int device;
  int need_x;  // x slot needs its own inverse transform
  bool have_transforms;
  float* d_rot;
  float* d_off;
  float* d_frot;  // fast-path rows (cellmath.h make_fast_row)
  // density grids: d_grids holds the last step's (mog_grids_ptr); the step
  // deposits into d_grids_next, which is all zero between steps, and its
  // gather clears the old d_grids, after which the two swap
  long long* d_grids;
  long long* d_grids_next;
  float2* d_S1;
  long long s1_elems;  // float2 elements of d_S1
  float2* d_S2;
  float4* d_force;
  float2* d_tw;
  float* d_gtab;
  float* d_kd;
  float2* d_conv;  // real-space kernel spectrum (solver convolution, compat physical)
  // closure features: deconvolved Green's function, world-frame force
  // grids, per-cell force and mass, per-cell interpolation blocks (8 float4)
  // (allocated by the first mog_closure_features call)
  float* d_gtab_dc;
  float4* d_force_dc;
  float4* d_fm;
  float4* d_blk;
  // FP64 diagnostics (allocated by the first mog_diagnostics call):
  // twiddles exp(-2 pi i j / g), the energy table T(k), reduction scratch
  double2* d_diag_tw;
  double* d_diag_tab;
  double* d_diag_scratch;
  // the spare density grids may hold a partial deposit (a step failed after
  // its deposit was enqueued): the next step clears them first
  bool spare_dirty;
  float* d_Wf;  // MFMA DFT tables (g >= 16): forward, inverse
  float* d_Wi;
  DepositTune tune;
  // Morton visiting order of the gather, rebuilt every `resort_every` steps
  // for the particle array it was built from
  int* d_order;
  int* d_idx;
  unsigned* d_keys_a;
  unsigned* d_keys_b;
  void* d_sort_tmp;
  size_t sort_tmp_bytes;
  long long order_cap;
  const void* order_pos;
  long long order_n;
  int steps_since_order;
  int resort_every;
  bool sorted_deposit;
  bool cell_cache;
  bool spectral_mfma;
  // sorted-step buffers (g <= 32): positions in visiting order and the
  // uint16 cell of every (grid, slot) pair
  float4* d_spos;
  unsigned short* d_cells;
  // prepared rows of the LDS deposit (centred coordinates, milli-mass)
  float4* d_cpos;
  long long cpos_cap;
  // mog_run's particle state in visiting order (the gather's reads and
  // writes at slot t instead of scattered over the caller's arrays)
  float4* d_ps;
  float4* d_vs;
  long long state_cap;
  const float4* g_ps;  // the state buffer the captured graphs use
  unsigned int g_prep;  // their gather writes the prepared rows (limit; 0: positions)

  // graph cache of mog_run: one captured step per assignment of the two
  // density buffers (gexec deposits into g_dep, gexec_alt into the other)
  hipStream_t cap_stream;
  hipGraphExec_t gexec;
  hipGraphExec_t gexec_alt;
  const long long* g_dep;
  const void* g_pos;
  const void* g_vel;
  long long g_n;
  bool g_ordered;
  const float4* g_cpos;
  std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(mog_ctx* ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(mog_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_create_error = buf;
  return code;
}

int hip_fail(mog_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, MOG_EHIP, "%s: %s", what, hipGetErrorString(e));
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// The reference's x force is overwritten by its z force when the number of
// radix-2 dispatches per 3D transform, 3*(log2 g + 1), is odd
// (main.py:564-590 returns the last destination; main.py:638-646 passes
// complex_tex_b as destination of both the x and the z transform).
bool aliases_x_with_z(int g) { return ((3 * (ilog2(g) + 1)) & 1) == 1; }

int validate(const mog_config* c) {
  if (!c) return fail(nullptr, MOG_EINVAL, "config is NULL");
  if (c->n_grids < 1 || c->n_grids > 65535)
    return fail(nullptr, MOG_EINVAL, "n_grids must be in [1, 65535], got %d", c->n_grids);
  if (!is_pow2(c->grid_size) || c->grid_size < 2 || c->grid_size > 64)
    return fail(nullptr, MOG_EINVAL, "grid_size must be a power of two in [2, 64], got %d",
                c->grid_size);
  if (!(isfinite(c->world_size) && c->world_size > 0.0))
    return fail(nullptr, MOG_EINVAL, "world_size must be finite and > 0");
  if (!isfinite(c->G)) return fail(nullptr, MOG_EINVAL, "G must be finite");
  if (!isfinite(c->dt)) return fail(nullptr, MOG_EINVAL, "dt must be finite");
  if (!isfinite(c->damping)) return fail(nullptr, MOG_EINVAL, "damping must be finite");
  if (!isfinite(c->timescale)) return fail(nullptr, MOG_EINVAL, "timescale must be finite");
  if (c->compat != MOG_COMPAT_REFERENCE && c->compat != MOG_COMPAT_PHYSICAL)
    return fail(nullptr, MOG_EINVAL, "compat must be 0 (reference) or 1 (physical)");
  if (c->precision < MOG_PRECISION_FP32 || c->precision > MOG_PRECISION_FP16)
    return fail(nullptr, MOG_EINVAL, "precision must be 0 (fp32), 1 (bf16x3) or 2 (fp16)");
  if (c->precision != MOG_PRECISION_FP32 && c->grid_size < 16)
    return fail(nullptr, MOG_EUNSUPPORTED,
                "precision %d (matrix-core DFT) needs grid_size >= 16, got %d", c->precision,
                c->grid_size);
  if (c->solver != MOG_SOLVER_FFT && c->solver != MOG_SOLVER_CONVOLUTION)
    return fail(nullptr, MOG_EINVAL, "solver must be 0 (fft) or 1 (convolution)");
  const double cells = (double)c->n_grids * c->grid_size * c->grid_size * c->grid_size;
  if (cells > 2147483647.0) return fail(nullptr, MOG_EINVAL, "n_grids * grid_size^3 exceeds 2^31");
  return MOG_OK;
}

// float32 ``k2 > 1e-4`` of greens.glsl:32-45 (PI = 3.14159265359f,
// k = float(shifted) * (TWO_PI / L), k2 = (kx*kx + ky*ky) + kz*kz).
bool green_mask32(int ix, int iy, int iz, int g, double W) {
  const float two_pi = 2.0f * 3.14159265359f;
  const float step = two_pi / (float)W;
  auto sh = [g](int i) { return i <= g / 2 ? i : i - g; };
  volatile float kx = (float)sh(ix) * step;
  volatile float ky = (float)sh(iy) * step;
  volatile float kz = (float)sh(iz) * step;
  volatile float a = kx * kx;
  volatile float b = ky * ky;
  volatile float c = kz * kz;
  volatile float k2 = a + b;
  k2 = k2 + c;
  return k2 > 1e-4f;
}

// DFT matrices for the matrix-core transforms, as the A-operand tables
// [Re W | -Im W] and [Im W | Re W] (g rows of 2g floats each), forward
// W = exp(-2 pi i jk/g) into Wf, inverse W = exp(+2 pi i jk/g) into Wi.
// 16-bit operands as the device converts them: BF16 and FP16 by
// round-to-nearest-even.
static uint16_t bf16_bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // finite inputs only (DFT coefficients)
  return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}
static uint16_t f16_bits(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t b;
  memcpy(&b, &h, 2);
  return b;
}

// Floats that follow a direction's FP32 tables: the A-operand fragments of
// its two row tables, for the 16x16x32 products 6 parts x (g/16)^2 tiles x
// 64 lanes x 8 values of 16 bits, then for the 16x16x4 FP32 products 2 parts
// x (g/16) row tiles x g/2 k-steps x 64 lanes (none below g = 16).
size_t dft_fragment_floats(int g) {
  if (g < 16) return 0;
  const size_t t = (size_t)(g / 16);
  return 6 * t * t * 64 * 8 / 2 + 2 * t * (size_t)(g / 2) * 64;
}

// Fragment (part p, row tile rt, k-block kb, lane l) = A[16 rt + l % 16]
// [32 kb + 8 (l / 16) + j], j < 8, of WA (parts 0, 1, 4) or WB (parts 2, 3,
// 5): BF16 hi and lo = bf16(a - hi) (parts 0-3), FP16 (parts 4, 5); the
// layout of one lane's operand register for the 16x16x32 products.
void build_dft_fragments(int g, const float* W, uint16_t* out) {
  if (g < 16) return;
  const int t = g / 16, g2 = 2 * g;
  const size_t nf = (size_t)t * t * 64;
  for (int rt = 0; rt < t; ++rt)
    for (int kb = 0; kb < t; ++kb)
      for (int l = 0; l < 64; ++l)
        for (int j = 0; j < 8; ++j) {
          const int row = 16 * rt + (l & 15), k = 32 * kb + 8 * (l >> 4) + j;
          const size_t at = (((size_t)rt * t + kb) * 64 + l) * 8 + j;
          const float a = W[row * g2 + k], b = W[2 * g * g + row * g2 + k];
          const uint16_t ah = bf16_bits(a), bh = bf16_bits(b);
          out[at] = ah;
          out[nf * 8 + at] = bf16_bits(a - bf16_value(ah));
          out[2 * nf * 8 + at] = bh;
          out[3 * nf * 8 + at] = bf16_bits(b - bf16_value(bh));
          out[4 * nf * 8 + at] = f16_bits(a);
          out[5 * nf * 8 + at] = f16_bits(b);
        }
  // FP32 parts for the 16x16x4 products: (row tile rt, k-step ks, lane l)
  // holds A[16 rt + l % 16][4 ks + l / 16]
  float* f32 = reinterpret_cast<float*>(out + 6 * nf * 8);
  const int nk = g / 2;
  const size_t nf32 = (size_t)t * nk * 64;
  for (int rt = 0; rt < t; ++rt)
    for (int ks = 0; ks < nk; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int row = 16 * rt + (l & 15), k = 4 * ks + (l >> 4);
        const size_t at = ((size_t)rt * nk + ks) * 64 + l;
        f32[at] = W[row * g2 + k];
        f32[nf32 + at] = W[2 * g * g + row * g2 + k];
      }
}

void build_dft_tables(int g, float* Wf, float* Wi) {
  const int g2 = 2 * g;
  for (int j = 0; j < g; ++j)
    for (int k = 0; k < g; ++k) {
      const int jk = (j * k) % g;
      const double th = 2.0 * M_PI * (double)jk / (double)g;
      const float cs = (float)cos(th), sn = (float)sin(th);
      // forward: Re W = cos, Im W = -sin
      Wf[j * g2 + k] = cs;
      Wf[j * g2 + g + k] = sn;
      Wf[2 * g * g + j * g2 + k] = -sn;
      Wf[2 * g * g + j * g2 + g + k] = cs;
      // inverse: Re W = cos, Im W = sin
      Wi[j * g2 + k] = cs;
      Wi[j * g2 + g + k] = -sn;
      Wi[2 * g * g + j * g2 + k] = sn;
      Wi[2 * g * g + j * g2 + g + k] = cs;
    }
}

void build_tables(const mog_config* c, float2* tw, float* gtab, float* kd) {
  const int g = c->grid_size;
  const double W = c->world_size;
  for (int j = 0; j < g; ++j) {
    const double th = 2.0 * M_PI * (double)j / (double)g;
    tw[j] = make_float2((float)cos(th), (float)sin(th));
  }
  auto sh = [g](int i) { return i <= g / 2 ? i : i - g; };
  for (int i = 0; i < g; ++i)
    kd[i] = (i == g / 2 && g > 1) ? 0.0f : (float)((double)sh(i) * 2.0 * M_PI / W);
  const double scale = 1.0 / (1000.0 * (double)g * g * g);
  for (int iz = 0; iz < g; ++iz)
    for (int iy = 0; iy < g; ++iy)
      for (int ix = 0; ix < g; ++ix) {
        double v = 0.0;
        if ((ix | iy | iz) != 0 && green_mask32(ix, iy, iz, g, W)) {
          const double kx = sh(ix) * 2.0 * M_PI / W, ky = sh(iy) * 2.0 * M_PI / W,
                       kz = sh(iz) * 2.0 * M_PI / W;
          const double k2 = kx * kx + ky * ky + kz * kz;
          v = -2.0 * M_PI * c->G / k2 * scale;
        }
        gtab[(iz * g + iy) * g + ix] = (float)v;
      }
}

// `solver: convolution` as a direct periodic real-space convolution with the
// Green's-gradient kernel of main.py:250-278: K_d(r) = r_d / (4 pi |r|^3) at
// the minimum-image cell offsets (index i -> i if i <= g/2 else i - g, the
// offset in units of voxel = W/g), K(0) = 0, values rounded to float32 as
// the reference stores them.  With rho the mass per cell, the force is
// F_d = 2 pi G voxel^3 (rho (*) K_d), the same coupling as the spectral
// solve (grad phi with del^2 phi = 2 pi G rho), so a single mass m gives
// G voxel^3 / 2 * m r / |r|^3 at every other cell.  The circular convolution
// is applied through the step's own transforms: the spectrum
// T_d = 2 pi G voxel^3 / (1000 g^3) * DFT(K_d) (FP64 separable DFT of the
// float32 kernel) multiplies the milli-mass spectrum, and the unnormalised
// inverse transforms give F exactly as they give the FFT solver's force.
// out[0..g^3) = T_y + i T_z, out[g^3..2 g^3) = T_x.
void build_conv_tables(const mog_config* c, float2* out) {
  const int g = c->grid_size;
  const size_t g3 = (size_t)g * g * g;
  const double voxel = c->world_size / (double)g;
  std::vector<std::complex<double>> K[3];
  for (int d = 0; d < 3; ++d) K[d].assign(g3, 0.0);
  auto sh = [g](int i) { return i <= g / 2 ? i : i - g; };
  for (int iz = 0; iz < g; ++iz)
    for (int iy = 0; iy < g; ++iy)
      for (int ix = 0; ix < g; ++ix) {
        const double rx = sh(ix) * voxel, ry = sh(iy) * voxel, rz = sh(iz) * voxel;
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double r = sqrt(r2);
        if (r > 1e-10) {
          const double f = 1.0 / (4.0 * M_PI * r2 * r);
          const size_t i = ((size_t)iz * g + iy) * g + ix;
          K[0][i] = (double)(float)(rx * f);
          K[1][i] = (double)(float)(ry * f);
          K[2][i] = (double)(float)(rz * f);
        }
      }
  // forward DFT (sign -) along x, y, z
  std::vector<std::complex<double>> w(g), line(g);
  for (int j = 0; j < g; ++j) w[j] = std::polar(1.0, -2.0 * M_PI * (double)j / (double)g);
  const size_t stride[3] = {1, (size_t)g, (size_t)g * g};
  for (int d = 0; d < 3; ++d)
    for (int axis = 0; axis < 3; ++axis) {
      const size_t st = stride[axis];
      for (size_t base = 0; base < g3; ++base) {
        if ((base / st) % g != 0) continue;  // one line per start index
        for (int kk = 0; kk < g; ++kk) {
          std::complex<double> acc = 0.0;
          for (int j = 0; j < g; ++j) acc += K[d][base + j * st] * w[(size_t)j * kk % g];
          line[kk] = acc;
        }
        for (int kk = 0; kk < g; ++kk) K[d][base + kk * st] = line[kk];
      }
    }
  const double scale = 2.0 * M_PI * c->G * voxel * voxel * voxel / (1000.0 * (double)g3);
  const std::complex<double> I(0.0, 1.0);
  for (size_t i = 0; i < g3; ++i) {
    const std::complex<double> yz = scale * (K[1][i] + I * K[2][i]), x = scale * K[0][i];
    out[i] = make_float2((float)yz.real(), (float)yz.imag());
    out[g3 + i] = make_float2((float)x.real(), (float)x.imag());
  }
  if (g == 32) {  // half-spectrum multipliers [d][kz][ky][kx <= g/2] of the g = 32 solve
    const int hx = g / 2 + 1;
    float2* half = out + 2 * g3;
    for (int d = 0; d < 3; ++d)
      for (int iz = 0; iz < g; ++iz)
        for (int iy = 0; iy < g; ++iy)
          for (int ix = 0; ix < hx; ++ix) {
            const std::complex<double> t = scale * K[d][((size_t)iz * g + iy) * g + ix];
            half[(((size_t)d * g + iz) * g + iy) * hx + ix] = make_float2((float)t.real(), (float)t.imag());
          }
  }
}

Params make_params(const mog_config* c) {
  Params P;
  P.k = c->n_grids;
  P.g = c->grid_size;
  P.lg = ilog2(c->grid_size);
  P.cells = c->grid_size * c->grid_size * c->grid_size;
  P.W = (float)c->world_size;
  P.half = P.W * 0.5f;
  P.h = (float)(c->world_size / (double)c->grid_size);
  P.dtp = (float)(c->dt * c->timescale);
  P.damping = (float)c->damping;
  P.fk = (float)c->n_grids;
  // Admission bound of the fast path: |c_j| <= W / 2 holds for every
  // position the integrator's wrap leaves in [0, W]; a particle outside takes
  // the exact sequence.  The error margin scales with this bound.
  P.CB = P.half;
  P.lim = -1.0f;  // set with the transforms
  P.shift_ok = 0;
  return P;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
}

void release(mog_ctx* c) {
  if (!c) return;
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->gexec_alt) (void)hipGraphExecDestroy(c->gexec_alt);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  (void)hipFree(c->d_rot);
  (void)hipFree(c->d_off);
  (void)hipFree(c->d_frot);
  (void)hipFree(c->d_grids);
  (void)hipFree(c->d_grids_next);
  (void)hipFree(c->d_S1);
  (void)hipFree(c->d_S2);
  (void)hipFree(c->d_force);
  (void)hipFree(c->d_tw);
  (void)hipFree(c->d_gtab);
  (void)hipFree(c->d_kd);
  (void)hipFree(c->d_conv);
  (void)hipFree(c->d_gtab_dc);
  (void)hipFree(c->d_force_dc);
  (void)hipFree(c->d_fm);
  (void)hipFree(c->d_blk);
  (void)hipFree(c->d_ps);
  (void)hipFree(c->d_vs);
  (void)hipFree(c->d_diag_tw);
  (void)hipFree(c->d_diag_tab);
  (void)hipFree(c->d_diag_scratch);
  (void)hipFree(c->d_order);
  (void)hipFree(c->d_idx);
  (void)hipFree(c->d_keys_a);
  (void)hipFree(c->d_keys_b);
  (void)hipFree(c->d_sort_tmp);
  (void)hipFree(c->d_Wf);
  (void)hipFree(c->d_Wi);
  (void)hipFree(c->d_cpos);

}

int check_particles(mog_ctx* ctx, const void* pos, const void* vel, int64_t n, bool need_vel) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (n < 0) return fail(ctx, MOG_EINVAL, "n must be >= 0, got %lld", (long long)n);
  if (n > (int64_t)2147483647) return fail(ctx, MOG_EINVAL, "n must be < 2^31, got %lld", (long long)n);
  if (n == 0) return MOG_OK;
  if (!pos) return fail(ctx, MOG_EINVAL, "pos is NULL");
  if (!aligned16(pos)) return fail(ctx, MOG_EINVAL, "pos must be 16-byte aligned");
  if (need_vel) {
    if (!vel) return fail(ctx, MOG_EINVAL, "vel is NULL");
    if (!aligned16(vel)) return fail(ctx, MOG_EINVAL, "vel must be 16-byte aligned");
  }
  return MOG_OK;
}

// Default deposit tiling (measured on MI355X, two-disk and Plummer
// workloads): about 32 KB of LDS counters per workgroup of 512 threads;
// when one grid alone needs more than 64 KB (g = 32) the CU holds a single
// workgroup, so it gets 1024 threads. About 4096 workgroups either way.
DepositTune default_tune(const Params& P) {
  DepositTune t;
  const size_t grid_bytes = (size_t)P.cells * 4;
  const int per = (int)((32 * 1024) / grid_bytes);
  t.grids_per_wg = per < 1 ? 1 : (per > 8 ? 8 : per);
  // measured optima on MI355X: g=16 (16 KB grids) 512 threads, 2 grids;
  // g=32 (128 KB) 1024 threads, 1 grid; 4096 workgroups for both
  t.threads = grid_bytes > 64 * 1024 ? 1024 : 512;
  t.wg_target = 4096;
  return t;
}

// Particle counts below this visit in storage order (sorting costs more than
// it saves).
const long long kOrderMin = 32768;

void free_order(mog_ctx* c) {
  (void)hipFree(c->d_spos);
  (void)hipFree(c->d_cells);
  c->d_spos = nullptr;
  c->d_cells = nullptr;
  (void)hipFree(c->d_order);
  (void)hipFree(c->d_idx);
  (void)hipFree(c->d_keys_a);
  (void)hipFree(c->d_keys_b);
  (void)hipFree(c->d_sort_tmp);
  c->d_order = c->d_idx = nullptr;
  c->d_keys_a = c->d_keys_b = nullptr;
  c->d_sort_tmp = nullptr;
  c->order_cap = 0;
  c->order_pos = nullptr;
}

// Buffers of the visiting order for n particles (outside graph capture).
// Returns false (no ordering) when n is small or allocation fails.
bool ensure_order(mog_ctx* c, int64_t n) {
  if (n < kOrderMin || c->resort_every <= 0) return false;
  if (c->order_cap >= n) return true;
  free_order(c);
  size_t tmp = 0;
  if (order_temp_bytes(n, &tmp) != hipSuccess) return false;
  const size_t b = (size_t)n * 4;
  if (hipMalloc((void**)&c->d_order, b) != hipSuccess || hipMalloc((void**)&c->d_idx, b) != hipSuccess ||
      hipMalloc((void**)&c->d_keys_a, b) != hipSuccess ||
      hipMalloc((void**)&c->d_keys_b, b) != hipSuccess ||
      hipMalloc(&c->d_sort_tmp, tmp > 0 ? tmp : 16) != hipSuccess) {
    (void)hipGetLastError();
    free_order(c);
    return false;
  }
  c->sort_tmp_bytes = tmp;
  c->order_cap = n;
  // sorted deposit + cached gather need the copies; without them the step
  // still uses the visiting order in the gather
  if (c->P.g <= 32 && c->sorted_deposit) {
    if (hipMalloc((void**)&c->d_spos, (size_t)n * sizeof(float4)) != hipSuccess ||
        hipMalloc((void**)&c->d_cells,
                  (size_t)((n + 63) / 64) * 64 * c->P.k * sizeof(unsigned short)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(c->d_spos);
      (void)hipFree(c->d_cells);
      c->d_spos = nullptr;
      c->d_cells = nullptr;
    }
  }
  return true;
}

// The deposit's prepared-row scratch for n particles (outside graph
// capture); NULL when allocation fails (the deposit then reads the
// positions directly).
float4* ensure_cpos(mog_ctx* c, int64_t n) {
  if (c->cpos_cap >= n && c->d_cpos) return c->d_cpos;
  (void)hipFree(c->d_cpos);
  c->d_cpos = nullptr;
  c->cpos_cap = 0;
  if (hipMalloc((void**)&c->d_cpos, (size_t)n * sizeof(float4)) != hipSuccess) {
    (void)hipGetLastError();
    c->d_cpos = nullptr;
    return nullptr;
  }
  c->cpos_cap = n;
  return c->d_cpos;
}

// mog_run's state buffers for n particles (outside graph capture); false
// when allocation fails (the run then gathers through the caller's arrays).
bool ensure_state(mog_ctx* c, int64_t n) {
  if (c->state_cap >= n && c->d_ps && c->d_vs) return true;
  (void)hipFree(c->d_ps);
  (void)hipFree(c->d_vs);
  c->d_ps = c->d_vs = nullptr;
  c->state_cap = 0;
  if (hipMalloc((void**)&c->d_ps, (size_t)n * sizeof(float4)) != hipSuccess ||
      hipMalloc((void**)&c->d_vs, (size_t)n * sizeof(float4)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(c->d_ps);
    (void)hipFree(c->d_vs);
    c->d_ps = c->d_vs = nullptr;
    return false;
  }
  c->state_cap = n;
  return true;
}

// Runs of at least this many steps keep their state in visiting order: the
// import and the export cost about two gathers' worth of scattered
// accesses, which a step through the copies saves about once.
constexpr int kStateMinSteps = 4;

// Whether refresh_order would rebuild the order for (pos, n).
bool order_stale(const mog_ctx* c, const float* pos, int64_t n) {
  return c->order_pos != pos || c->order_n != n || c->steps_since_order >= c->resort_every;
}

// Rebuild the order when it is stale for (pos, n); returns the order to use
// (NULL: storage order).
hipError_t refresh_order(mog_ctx* c, const float* pos, int64_t n, hipStream_t s, const int** out) {
  *out = nullptr;
  if (!ensure_order(c, n)) return hipSuccess;
  if (c->order_pos != pos || c->order_n != n || c->steps_since_order >= c->resort_every) {
    hipError_t e = build_order((const float4*)pos, n, c->P.W, c->d_keys_a, c->d_keys_b, c->d_idx,
                               c->d_order, c->d_sort_tmp, c->sort_tmp_bytes, s);
    if (e != hipSuccess) return e;
    c->order_pos = pos;
    c->order_n = n;
    c->steps_since_order = 0;
  }
  c->steps_since_order++;
  *out = c->d_order;
  return hipSuccess;
}

// Spectral solve: matrix-core passes for g >= 16, vector passes below.
hipError_t solve_dispatch(mog_ctx* ctx, const long long* grids, float4* force, int world,
                          hipStream_t s) {
  const Params& P = ctx->P;
  if (P.g >= 16 && (ctx->spectral_mfma || ctx->cfg.precision != MOG_PRECISION_FP32))
    return launch_solve_mfma(P, ctx->cfg.precision, grids, ctx->d_S1, ctx->d_S2, force, ctx->d_Wf,
                             ctx->d_Wi, ctx->d_gtab, ctx->d_kd, ctx->d_conv, ctx->need_x,
                             ctx->d_rot, world, s);
  return launch_solve(P, grids, ctx->d_S1, ctx->s1_elems, ctx->d_S2, force, ctx->d_tw,
                      ctx->d_gtab, ctx->d_kd,
                      ctx->d_conv, ctx->need_x, ctx->d_rot, world, s);
}

void drop_graphs(mog_ctx* ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->gexec_alt) (void)hipGraphExecDestroy(ctx->gexec_alt);
  ctx->gexec = ctx->gexec_alt = nullptr;
}

// One step on stream s: deposit into `dep` (all zero), solve, gather +
// update, and clear `old` on the way (the gather's side job).  The caller
// then makes dep the current density and old the zero spare (swap_grids)
// once the step is launched, not when it is only captured.
hipError_t enqueue_step(mog_ctx* ctx, float* pos, float* vel, int64_t n, const int* order,
                        long long* dep, long long* old, hipStream_t s, bool allow_sorted = true,
                        float4* cpos = nullptr, float4* ps = nullptr, float4* vs = nullptr,
                        unsigned int* gather_prep = nullptr) {
  // gather_prep (with ps, vs): when the deposit reads prepared rows, it runs
  // no prepass and the gather writes the next step's rows instead of the
  // positions; *gather_prep is set to their milli-mass limit (0: not so)
  if (gather_prep) *gather_prep = 0;
  const Params& P = ctx->P;
  const long long cells = (long long)P.k * P.cells;
  if (allow_sorted && order && ctx->sorted_deposit && ctx->d_spos && ctx->d_cells) {
    // sorted step: positions in visiting order, deposit with aggregated
    // atomics (and a cell cache), gather from the cache or recomputed cells
    const bool cache = ctx->cell_cache;
    hipError_t e = launch_permute((const float4*)pos, order, n, ctx->d_spos, s);
    if (e == hipSuccess)
      e = launch_deposit_ex(P, ctx->d_spos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                            (unsigned long long*)dep, cache ? ctx->d_cells : nullptr, true,
                            ctx->tune, s);
    if (e == hipSuccess) e = solve_dispatch(ctx, dep, ctx->d_force, 1, s);
    if (e == hipSuccess) {
      if (cache) {
        e = launch_gather_cached(P, (float4*)pos, (float4*)vel, n, order, ctx->d_cells,
                                 ctx->d_force, s);
        if (e == hipSuccess) e = launch_fill_zero(old, cells, s);
      } else {
        e = launch_gather_update(P, (float4*)pos, (float4*)vel, n, ctx->d_rot, ctx->d_off,
                                 ctx->d_frot, ctx->d_force, 1, nullptr, 0, order, s, old, cells);
      }
    }
    return e;
  }
  const bool prep_by_gather = gather_prep && order && ps && vs && cpos;
  DepositTune dt = ctx->tune;
  dt.cpos_ready = prep_by_gather;
  unsigned int lim = 0;
  hipError_t e = launch_deposit(P, (const float4*)pos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                                (unsigned long long*)dep, dt, s, cpos, &lim);
  if (e != hipSuccess) return e;
  e = solve_dispatch(ctx, dep, ctx->d_force, 1, s);
  if (e != hipSuccess) return e;
  const bool prepped = prep_by_gather && lim != 0;
  if (prepped) *gather_prep = lim;
  return launch_gather_update(P, (float4*)pos, (float4*)vel, n, ctx->d_rot, ctx->d_off,
                              ctx->d_frot, ctx->d_force, 1, nullptr, 0, order, s, old, cells,
                              order ? ps : nullptr, order ? vs : nullptr,
                              prepped ? cpos : nullptr, prepped ? lim : 0u);
}

// Which visiting order the no-allocation scratch of each device holds
// (mog_step_noalloc): the context, particle array and count it was built
// for, the steps since, and the index buffer holding it.
struct NoAllocOrder {
  unsigned long long owner;
  const void* pos;
  long long n;
  int steps;
  const int* order;
};
std::mutex g_noalloc_mu;
NoAllocOrder g_noalloc_order[64] = {};

void swap_grids(mog_ctx* ctx) {
  long long* t = ctx->d_grids;
  ctx->d_grids = ctx->d_grids_next;
  ctx->d_grids_next = t;
}

// A step that fails after its deposit was enqueued leaves part of a
// density in the spare grids (no swap follows); the next step clears them
// before it deposits into them.
hipError_t clean_spare(mog_ctx* ctx, hipStream_t s) {
  if (!ctx->spare_dirty) return hipSuccess;
  hipError_t e = launch_fill_zero(ctx->d_grids_next, (long long)ctx->P.k * ctx->P.cells, s);
  if (e == hipSuccess) ctx->spare_dirty = false;
  return e;
}


// ---------------------------------------------------------------------------
// Batched ensemble
// ---------------------------------------------------------------------------
}  // namespace

// grid sizes 2, 4, 8, 16, 32 each form a solve class of the batch
constexpr int kBatchClasses = 5;

struct mog_batch {
  int nsims;
  std::vector<mog_config> cfg;
  std::vector<BatchSim> sims;  // host copy of the device descriptor table
  std::vector<char> have_transforms;
  long long ntotal, cells_total, gather_blocks;
  int grids_total;
  int kmax;  // largest n_grids of the simulations
  // solve classes, one per grid size g = 2 .. 32 (class log2 g - 1):
  // simulations, slabs, LDS
  struct SolveClass {
    int start, count, gmax;
    long long slabs;
  } cls[kBatchClasses];
  int* d_cls_sim;
  long long* d_cls_off;
  float4* d_cpos;  // the deposit's prepared rows, one per particle
  // solve windows (kBatchWindowCells): class, first slab and slab count,
  // and per window a row of n_sims scratch offsets (class slot j -> offset
  // of the simulation's grid 0 in the window's scratch)
  struct SolveWindow {
    int cls;
    long long slab0, slabs;
  };
  std::vector<SolveWindow> windows;
  long long* d_win_soff;
  BatchSim* d_sims;
  float* d_rot;
  float* d_off;
  float* d_frot;
  long long* d_grids;
  float2* d_S1;
  long long s1_elems;
  float2* d_S2;
  float4* d_force;
  float2* d_tw;
  float* d_gtab;
  float* d_kd;
  float2* d_conv;
  DepositTask* d_tasks;
  int ntasks;
  unsigned long long* d_counter;
  // visiting order of the gather (Morton order inside each simulation),
  // rebuilt every `resort_every` steps outside any captured graph
  int* d_order;
  int* d_idx;
  unsigned long long* d_keys_a;
  unsigned long long* d_keys_b;
  void* d_sort_tmp;
  size_t sort_tmp_bytes;
  const void* order_pos;
  int steps_since_order;
  int resort_every;
  // mog_batch_run's state in visiting order (allocated by its first run)
  float4* d_ps;
  float4* d_vs;
  long long deposit_chunk;
  int deposit_threads;
  hipStream_t cap_stream;
  hipGraphExec_t gexec;
  const void* g_pos;
  const void* g_vel;
  const void* g_ps;  // the state buffer the captured graph uses
};

struct mog2d_ctx {
  int k, g, lg;
  double W, G;
  float lim;  // fast-path acceptance 0.5 - margin (< 0: exact sequence only)
  float* d_rot;
  float* d_off;
  float* d_frot;
  float2* d_tw;
  float* d_gtab;
  float* d_kd;
  long long* d_grids;
  float2* d_force;
  float4* d_avg;
  long long avg_cap;
};

namespace {

void mog2d_release(mog2d_ctx* c) {
  void* ptrs[] = {c->d_rot, c->d_off, c->d_frot, c->d_tw, c->d_gtab, c->d_kd, c->d_grids, c->d_force,
                  c->d_avg};
  for (void* q : ptrs) (void)hipFree(q);
}

void batch_release(mog_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  void* ptrs[] = {b->d_sims, b->d_rot,  b->d_off, b->d_frot, b->d_grids, b->d_S1,
                  b->d_S2,   b->d_force, b->d_tw, b->d_gtab, b->d_kd,    b->d_conv,
                  b->d_tasks, b->d_counter, b->d_order, b->d_idx, b->d_keys_a, b->d_keys_b,
                  b->d_sort_tmp, b->d_cls_sim, b->d_cls_off, b->d_win_soff, b->d_cpos,
                  b->d_ps, b->d_vs};
  for (void* q : ptrs) (void)hipFree(q);
}

void batch_drop_graph(mog_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
}

// Deposit tasks: every simulation is cut into particle chunks of
// `chunk` x groups of grids whose uint32 histograms fit kBatchLdsCells;
// ordered by cost (pairs + histogram cells), largest first.
hipError_t batch_build_tasks(mog_batch* b) {
  std::vector<DepositTask> tasks;
  std::vector<double> cost;
  for (int si = 0; si < b->nsims; ++si) {
    const BatchSim& sm = b->sims[si];
    if (sm.n == 0) continue;
    const int cells = sm.P.cells;
    const int per = batch_task_grids(cells);
    for (long long p0 = 0; p0 < sm.n; p0 += b->deposit_chunk)
      for (int g0 = 0; g0 < sm.P.k; g0 += per) {
        DepositTask t;
        t.sim = si;
        t.grid0 = g0;
        t.ng = std::min(per, sm.P.k - g0);
        t.G = per;
        t.p0 = p0;
        t.p1 = std::min(sm.n, p0 + b->deposit_chunk);
        tasks.push_back(t);
        cost.push_back((double)(t.p1 - t.p0) * t.ng + 0.25 * (double)t.ng * cells);
      }
  }
  std::vector<size_t> idx(tasks.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
  std::vector<DepositTask> sorted(tasks.size());
  for (size_t i = 0; i < idx.size(); ++i) sorted[i] = tasks[idx[i]];
  (void)hipFree(b->d_tasks);
  b->d_tasks = nullptr;
  b->ntasks = (int)sorted.size();
  hipError_t e = dalloc(&b->d_tasks, sorted.size());
  if (e == hipSuccess && !sorted.empty())
    e = hipMemcpy(b->d_tasks, sorted.data(), sorted.size() * sizeof(DepositTask),
                  hipMemcpyHostToDevice);
  return e;
}

// Rebuild the visiting order when it is stale for `pos` (direct launches,
// never inside a capture).
hipError_t batch_refresh_order(mog_batch* b, const float* pos, hipStream_t s) {
  if (!b->d_order || b->resort_every <= 0) return hipSuccess;
  if (b->order_pos != pos || b->steps_since_order >= b->resort_every) {
    hipError_t e = build_batch_order((const float4*)pos, b->ntotal, b->gather_blocks, b->d_sims,
                                     b->nsims, b->d_keys_a, b->d_keys_b, b->d_idx, b->d_order,
                                     b->d_sort_tmp, b->sort_tmp_bytes, s);
    if (e != hipSuccess) return e;
    b->order_pos = pos;
    b->steps_since_order = 0;
  }
  b->steps_since_order++;
  return hipSuccess;
}

const int* batch_order(const mog_batch* b) {
  return (b->d_order && b->resort_every > 0) ? b->d_order : nullptr;
}

hipError_t batch_enqueue(mog_batch* b, float* pos, float* vel, hipStream_t s,
                         float4* ps = nullptr, float4* vs = nullptr) {
  hipError_t e = launch_fill_zero(b->d_grids, b->cells_total, s);
  if (e == hipSuccess) e = launch_fill_zero((long long*)b->d_counter, 1, s);
  if (e == hipSuccess)
    e = launch_batch_deposit((const float4*)pos, b->d_cpos, b->d_sims, b->nsims, b->gather_blocks,
                             b->deposit_chunk, b->d_tasks, b->ntasks, b->d_rot, b->d_off,
                             b->d_frot, (unsigned long long*)b->d_grids, b->d_counter,
                             b->deposit_threads, s);
  for (size_t w = 0; w < b->windows.size() && e == hipSuccess; ++w) {
    const mog_batch::SolveWindow& win = b->windows[w];
    const mog_batch::SolveClass& sc = b->cls[win.cls];
    e = launch_batch_solve(b->d_sims, b->d_cls_sim + sc.start, b->d_cls_off + sc.start, sc.count,
                           win.slab0, win.slabs, b->d_win_soff + w * (size_t)b->nsims, sc.gmax,
                           b->d_grids, b->d_S1, b->s1_elems, b->d_S2, b->d_force, b->d_tw,
                           b->d_gtab, b->d_kd, b->d_conv, b->d_rot, s);
  }
  for (int ci = 0; ci < kBatchClasses && e == hipSuccess; ++ci) {
    const mog_batch::SolveClass& sc = b->cls[ci];
    if (sc.count == 0 || !batch_class_fused(sc.gmax)) continue;
    e = launch_batch_solve(b->d_sims, b->d_cls_sim + sc.start, b->d_cls_off + sc.start, sc.count,
                           0, sc.slabs, nullptr, sc.gmax, b->d_grids, b->d_S1, b->s1_elems,
                           b->d_S2, b->d_force, b->d_tw, b->d_gtab, b->d_kd, b->d_conv, b->d_rot,
                           s);
  }
  if (e == hipSuccess)
    e = launch_batch_gather((float4*)pos, (float4*)vel, b->gather_blocks, b->d_sims, b->nsims,
                            b->kmax, b->d_rot, b->d_off, b->d_frot, b->d_force, batch_order(b), s,
                            batch_order(b) ? ps : nullptr, batch_order(b) ? vs : nullptr);
  return e;
}

// mog_batch_run's state buffers (outside capture); false when allocation
// fails (the run then gathers through the caller's arrays).
bool batch_ensure_state(mog_batch* b) {
  if (b->d_ps && b->d_vs) return true;
  if (hipMalloc((void**)&b->d_ps, (size_t)b->ntotal * sizeof(float4)) != hipSuccess ||
      hipMalloc((void**)&b->d_vs, (size_t)b->ntotal * sizeof(float4)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(b->d_ps);
    (void)hipFree(b->d_vs);
    b->d_ps = b->d_vs = nullptr;
    return false;
  }
  return true;
}

int batch_check(mog_batch* b, const float* pos, const float* vel) {
  if (!b) return fail(nullptr, MOG_EINVAL, "batch is NULL");
  if (b->ntotal > 0 && (!pos || !vel || !aligned16(pos) || !aligned16(vel)))
    return fail(nullptr, MOG_EINVAL, "pos/vel must be 16-byte aligned device pointers");
  for (int i = 0; i < b->nsims; ++i)
    if (!b->have_transforms[i])
      return fail(nullptr, MOG_ENOTRANSFORMS, "mog_batch_set_transforms not called for sim %d", i);
  return MOG_OK;
}

}  // namespace

extern "C" {

int mog_abi_version(void) { return MOG_ABI_VERSION; }

int mog_create(const mog_config* cfg, mog_ctx** out) {
  if (!out) return fail(nullptr, MOG_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc != MOG_OK) return rc;
  static std::atomic<unsigned long long> next_uid{1};
  mog_ctx* c = new mog_ctx();
  c->uid = next_uid.fetch_add(1);
  c->cfg = *cfg;
  c->P = make_params(cfg);
  c->need_x = !(cfg->compat == MOG_COMPAT_REFERENCE && aliases_x_with_z(cfg->grid_size));
  c->tune = default_tune(c->P);
  c->resort_every = 16;
  c->sorted_deposit = false;
  c->cell_cache = false;
  // the FP32 vector transforms beat the FP32 matrix-core ones at g = 16 and
  // 32 (the FP32 MFMA rate equals the packed vector rate); the split
  // precisions always use the matrix cores
  c->spectral_mfma = false;

  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    delete c;
    return fail(nullptr, MOG_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  }
  const size_t k = (size_t)cfg->n_grids, g = (size_t)cfg->grid_size, cells = k * g * g * g;
  // the matrix-core passes use S1 as one complex field; the g = 32 vector
  // solve as one scratch slot per workgroup
  c->s1_elems = (long long)cells;
  if (g == 32) c->s1_elems = std::max(c->s1_elems, rh_scratch_elems((long long)k));
  if (dalloc(&c->d_rot, k * 12) != hipSuccess || dalloc(&c->d_off, k * 4) != hipSuccess ||
      dalloc(&c->d_frot, k * 12) != hipSuccess ||
      dalloc(&c->d_grids, cells) != hipSuccess || dalloc(&c->d_grids_next, cells) != hipSuccess ||
      dalloc(&c->d_S1, (size_t)c->s1_elems) != hipSuccess ||
      dalloc(&c->d_S2, c->need_x ? cells : 1) != hipSuccess ||
      dalloc(&c->d_force, cells) != hipSuccess || dalloc(&c->d_tw, g) != hipSuccess ||
      dalloc(&c->d_gtab, g * g * g) != hipSuccess || dalloc(&c->d_kd, g) != hipSuccess ||
      dalloc(&c->d_Wf, 4 * g * g + dft_fragment_floats((int)g)) != hipSuccess ||
      dalloc(&c->d_Wi, 4 * g * g + dft_fragment_floats((int)g)) != hipSuccess) {
    release(c);
    delete c;
    return fail(nullptr, MOG_ENOMEM, "device allocation failed for k=%d g=%d", cfg->n_grids,
                cfg->grid_size);
  }
  float2* tw = (float2*)malloc(g * sizeof(float2));
  float* gtab = (float*)malloc(g * g * g * sizeof(float));
  float* kd = (float*)malloc(g * sizeof(float));
  const size_t wsz = 4 * g * g + dft_fragment_floats((int)g);
  float* Wf = (float*)malloc(wsz * sizeof(float));
  float* Wi = (float*)malloc(wsz * sizeof(float));
  build_tables(cfg, tw, gtab, kd);
  build_dft_tables((int)g, Wf, Wi);
  build_dft_fragments((int)g, Wf, (uint16_t*)(Wf + 4 * g * g));
  build_dft_fragments((int)g, Wi, (uint16_t*)(Wi + 4 * g * g));
  e = hipMemcpy(c->d_tw, tw, g * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_gtab, gtab, g * g * g * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_kd, kd, g * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_Wf, Wf, wsz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_Wi, Wi, wsz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->d_grids, 0, cells * sizeof(long long));
  if (e == hipSuccess) e = hipMemset(c->d_grids_next, 0, cells * sizeof(long long));
  if (e == hipSuccess && cfg->solver == MOG_SOLVER_CONVOLUTION &&
      cfg->compat == MOG_COMPAT_PHYSICAL) {
    // the reference runs the FFT solver for either key (main.py:671-672);
    // the real-space convolution is the physical mode's meaning of the key
    std::vector<float2> conv((size_t)conv_table_elems((int)g));
    build_conv_tables(cfg, conv.data());
    if (dalloc(&c->d_conv, conv.size()) != hipSuccess) e = hipErrorOutOfMemory;
    if (e == hipSuccess)
      e = hipMemcpy(c->d_conv, conv.data(), conv.size() * sizeof(float2), hipMemcpyHostToDevice);
  }
  free(Wf);
  free(Wi);
  free(tw);
  free(gtab);
  free(kd);
  if (e != hipSuccess) {
    release(c);
    delete c;
    return fail(nullptr, MOG_EHIP, "table upload: %s", hipGetErrorString(e));
  }
  *out = c;
  return MOG_OK;
}

void mog_destroy(mog_ctx* ctx) {
  if (!ctx) return;
  release(ctx);
  delete ctx;
}

int mog_set_transforms(mog_ctx* ctx, const float* rot_k12, const float* off_k4) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (!rot_k12 || !off_k4) return fail(ctx, MOG_EINVAL, "rot/off is NULL");
  const size_t k = (size_t)ctx->P.k;
  Params& P = ctx->P;
  double rmax = 0.0, omax = 0.0;
  bool finite = true;
  for (size_t i = 0; i < k * 12; ++i) {
    finite = finite && isfinite(rot_k12[i]);
    rmax = fmax(rmax, fabs((double)rot_k12[i]));
  }
  for (size_t i = 0; i < k; ++i)
    for (int m = 0; m < 3; ++m) {
      finite = finite && isfinite(off_k4[4 * i + m]);
      omax = fmax(omax, fabs((double)off_k4[4 * i + m]));
    }
  if (!finite) return fail(ctx, MOG_EINVAL, "rotations/offsets must be finite");
  float* frot = (float*)malloc(k * 12 * sizeof(float));
  if (!frot) return fail(ctx, MOG_ENOMEM, "host allocation failed");
  for (size_t i = 0; i < k; ++i) make_fast_row(rot_k12 + 12 * i, off_k4 + 4 * i, P.half, P.h, frot + 12 * i);
  const double margin = fast_margin(ctx->cfg.world_size, P.g, rmax, omax, P.CB);
  hipError_t e = hipMemcpy(ctx->d_rot, rot_k12, k * 12 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ctx->d_off, off_k4, k * 4 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ctx->d_frot, frot, k * 12 * sizeof(float), hipMemcpyHostToDevice);
  free(frot);
  if (e != hipSuccess) return hip_fail(ctx, e, "mog_set_transforms");
  // The fast path is only admitted with a small margin (sane transforms).
  P.lim = margin < 0.05 ? (float)(0.5 - margin) : -1.0f;
  P.shift_ok = shift_flags(ctx->cfg.world_size, P.g, P.lg, rmax, omax, P.CB);
  drop_graphs(ctx);  // kernel arguments of a captured step changed
  ctx->have_transforms = true;
  return MOG_OK;
}

int mog_step(mog_ctx* ctx, float* pos, float* vel, int64_t n, void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (n == 0) return MOG_OK;
  const int* order = nullptr;
  hipError_t e = clean_spare(ctx, (hipStream_t)stream);
  if (e == hipSuccess) e = refresh_order(ctx, pos, n, (hipStream_t)stream, &order);
  float4* const cpos = ensure_cpos(ctx, n);
  if (e == hipSuccess)
    e = enqueue_step(ctx, pos, vel, n, order, ctx->d_grids_next, ctx->d_grids, (hipStream_t)stream,
                     true, cpos);
  if (e != hipSuccess) {
    ctx->spare_dirty = true;
    return hip_fail(ctx, e, "mog_step");
  }
  swap_grids(ctx);
  return MOG_OK;
}

int mog_run(mog_ctx* ctx, float* pos, float* vel, int64_t n, int32_t nsteps, int32_t use_graph,
            void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (nsteps < 0) return fail(ctx, MOG_EINVAL, "nsteps must be >= 0");
  if (n == 0 || nsteps == 0) return MOG_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = clean_spare(ctx, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "mog_run");
  float4* const cpos = ensure_cpos(ctx, n);
  // With a visiting order, the plain deposit and at least kStateMinSteps
  // steps, the run keeps the particle state in visiting order (d_ps, d_vs):
  // imported at the start and after
  // each rebuild of the order, the gather reads and writes it at slot t and
  // writes the positions back to the caller's array (the deposit's input);
  // the velocities go back at each rebuild and at the end.  The step is the
  // same arithmetic on the same values, so the results are those of the
  // gather through the caller's arrays.
  const int* const order_buf_before = ctx->d_order;  // ensure_order may reallocate it
  const bool ordered = ensure_order(ctx, n);
  const bool state = nsteps >= kStateMinSteps && ordered &&
                     !(ctx->sorted_deposit && ctx->d_spos && ctx->d_cells) && ensure_state(ctx, n);
  float4* const ps = state ? ctx->d_ps : nullptr;
  float4* const vs = state ? ctx->d_vs : nullptr;
  bool imported = false;
  // graph replays whose gather writes the prepared rows (prep != 0) leave
  // pos stale as well until the export, and need the rows of the caller's
  // positions before the first replay
  unsigned int prep = 0;
  // the order for step i (rebuilt when stale), with the state moved around
  // the rebuild
  auto prepare = [&](int i, const int** order) -> hipError_t {
    *order = nullptr;
    if (!ordered) return hipSuccess;
    hipError_t ee = hipSuccess;
    const bool rebuild = order_stale(ctx, pos, n);
    if (state && imported && rebuild) {
      ee = launch_state_export(vs, ctx->d_order, n, (float4*)vel, s, prep ? ps : nullptr,
                               prep ? (float4*)pos : nullptr);
      if (ee != hipSuccess) return ee;
      imported = false;
    }
    if ((ee = refresh_order(ctx, pos, n, s, order)) != hipSuccess) return ee;
    if (state && (i == 0 || rebuild)) {
      if ((ee = launch_state_import((const float4*)pos, (const float4*)vel, *order, n, ps, vs, s)) !=
          hipSuccess)
        return ee;
      if (i == 0 && prep &&
          (ee = launch_deposit_prep(ctx->P, (const float4*)pos, n, prep, cpos, s)) != hipSuccess)
        return ee;
      imported = true;
    }
    return hipSuccess;
  };
  auto finish = [&](hipError_t ee, const char* what) -> int {
    if (state && imported) {
      const hipError_t ex = launch_state_export(vs, ctx->d_order, n, (float4*)vel, s,
                                                prep ? ps : nullptr, prep ? (float4*)pos : nullptr);
      if (ee == hipSuccess) ee = ex;
    }
    if (ee != hipSuccess) {
      ctx->spare_dirty = true;
      return hip_fail(ctx, ee, what);
    }
    return MOG_OK;
  };
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i) {
      const int* order = nullptr;
      if ((e = prepare(i, &order)) != hipSuccess ||
          (e = enqueue_step(ctx, pos, vel, n, order, ctx->d_grids_next, ctx->d_grids, s, true,
                            cpos, ps, vs)) != hipSuccess)
        return finish(e, "mog_run");
      swap_grids(ctx);
    }
    return finish(hipSuccess, "mog_run");
  }
  // The captured step reads ctx->d_order; the order itself is rebuilt by
  // direct launches between graph replays (the sort is not captured).
  if (ctx->gexec && (ctx->d_order != order_buf_before || ordered != ctx->g_ordered ||
                     ctx->g_cpos != cpos || ctx->g_ps != ps))
    drop_graphs(ctx);
  if (!ctx->gexec || ctx->g_pos != pos || ctx->g_vel != vel || ctx->g_n != n) {
    drop_graphs(ctx);
    if (!ctx->cap_stream &&
        (e = hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(ctx, e, "hipStreamCreate");
    // both assignments of the density buffers: a replay deposits into the
    // current spare and clears the current density
    long long* const bufs[2] = {ctx->d_grids_next, ctx->d_grids};
    hipGraphExec_t* const slot[2] = {&ctx->gexec, &ctx->gexec_alt};
    for (int a = 0; a < 2; ++a) {
      if ((e = hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal)) !=
          hipSuccess) {
        drop_graphs(ctx);
        return hip_fail(ctx, e, "hipStreamBeginCapture");
      }
      unsigned int gp = 0;
      hipError_t ee = enqueue_step(ctx, pos, vel, n, ordered ? ctx->d_order : nullptr, bufs[a],
                                   bufs[1 - a], ctx->cap_stream, true, cpos, ps, vs,
                                   state ? &gp : nullptr);
      ctx->g_prep = gp;
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(ctx->cap_stream, &graph);
      if (ee != hipSuccess || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        drop_graphs(ctx);
        return hip_fail(ctx, ee != hipSuccess ? ee : e, "capture");
      }
      e = hipGraphInstantiate(slot[a], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        *slot[a] = nullptr;
        drop_graphs(ctx);
        return hip_fail(ctx, e, "hipGraphInstantiate");
      }
    }
    ctx->g_dep = bufs[0];
    ctx->g_pos = pos;
    ctx->g_vel = vel;
    ctx->g_n = n;
    ctx->g_ordered = ordered;
    ctx->g_cpos = cpos;
    ctx->g_ps = ps;
  }
  prep = state ? ctx->g_prep : 0;
  for (int i = 0; i < nsteps; ++i) {
    const int* order = nullptr;
    if ((e = prepare(i, &order)) != hipSuccess) return finish(e, "mog_run");
    hipGraphExec_t gx = ctx->d_grids_next == ctx->g_dep ? ctx->gexec : ctx->gexec_alt;
    if ((e = hipGraphLaunch(gx, s)) != hipSuccess) return finish(e, "hipGraphLaunch");
    swap_grids(ctx);
  }
  return finish(hipSuccess, "mog_run");
}

int mog_step_noalloc(mog_ctx* ctx, float* pos, float* vel, int64_t n, void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (n > kNoAllocMaxParticles)
    return fail(ctx, MOG_EINVAL,
                "mog_step_noalloc: n = %lld exceeds its maximum of %lld particles (use mog_step)",
                (long long)n, kNoAllocMaxParticles);
  if (n == 0) return MOG_OK;
  hipStream_t s = (hipStream_t)stream;
  void* scr = nullptr;
  hipError_t e = noalloc_scratch(&scr);
  if (e != hipSuccess) return hip_fail(ctx, e, "mog_step_noalloc: scratch lookup");
  const int* order = nullptr;
  if (n >= kOrderMin && ctx->resort_every > 0) {
    // four n-element arrays of the visiting order, then the sort's
    // temporaries; the order kept there is reused while it belongs to this
    // context, particle array and count and is fresh enough
    const long long cap = kNoAllocMaxParticles;
    unsigned* keys_a = (unsigned*)scr;
    unsigned* keys_b = keys_a + cap;
    int* idx_a = (int*)(keys_b + cap);
    int* idx_b = idx_a + cap;
    void* tmp = (void*)(idx_b + cap);
    size_t need = 0;
    if ((e = order_temp_bytes_db(n, &need)) != hipSuccess)
      return hip_fail(ctx, e, "mog_step_noalloc: sort size");
    if (need > kNoAllocSortBytes)
      return fail(ctx, MOG_EUNSUPPORTED,
                  "mog_step_noalloc: the radix sort needs %zu bytes of temporaries, over the %zu "
                  "the scratch holds",
                  need, kNoAllocSortBytes);
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(ctx, e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocOrder& rec = g_noalloc_order[dev & 63];
    if (rec.owner != ctx->uid || rec.pos != pos || rec.n != n || !rec.order ||
        rec.steps >= ctx->resort_every) {
      rec.owner = 0;
      if ((e = build_order_db((const float4*)pos, n, ctx->P.W, keys_a, keys_b, idx_a, idx_b, tmp,
                              kNoAllocSortBytes, &rec.order, s)) != hipSuccess)
        return hip_fail(ctx, e, "mog_step_noalloc: visiting order");
      rec.owner = ctx->uid;
      rec.pos = pos;
      rec.n = n;
      rec.steps = 0;
    }
    rec.steps++;
    order = rec.order;
  }
  if ((e = clean_spare(ctx, s)) == hipSuccess)
    e = enqueue_step(ctx, pos, vel, n, order, ctx->d_grids_next, ctx->d_grids, s, false);
  if (e != hipSuccess) {
    ctx->spare_dirty = true;
    return hip_fail(ctx, e, "mog_step_noalloc");
  }
  swap_grids(ctx);
  return MOG_OK;
}

int mog_step_profiled(mog_ctx* ctx, float* pos, float* vel, int64_t n, float* phase_ms,
                      void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (!phase_ms) return fail(ctx, MOG_EINVAL, "phase_ms is NULL");
  hipStream_t s = (hipStream_t)stream;
  const Params& P = ctx->P;
  const int* order = nullptr;
  hipError_t eo = clean_spare(ctx, s);
  if (eo == hipSuccess) eo = refresh_order(ctx, pos, n, s, &order);
  if (eo != hipSuccess) return hip_fail(ctx, eo, "mog_step_profiled");
  hipEvent_t ev[5] = {};
  hipError_t e = hipSuccess;
  const bool sorted = order && ctx->sorted_deposit && ctx->d_spos && ctx->d_cells;
  const bool cache = sorted && ctx->cell_cache;
  // the gather of mog_run (state in visiting order), imported and exported
  // outside the timed phases
  const bool state = order && !sorted && ensure_state(ctx, n);
  float4* const ps = state ? ctx->d_ps : nullptr;
  float4* const vs = state ? ctx->d_vs : nullptr;
  for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  if (e == hipSuccess && state)
    e = launch_state_import((const float4*)pos, (const float4*)vel, order, n, ps, vs, s);
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  // the step's own sequence (its clear phase is empty: the grids it deposits
  // into were cleared by the previous step's gather)
  long long* const dep = ctx->d_grids_next;
  long long* const old = ctx->d_grids;
  const long long cells = (long long)P.k * P.cells;
  if (e == hipSuccess && sorted) e = launch_permute((const float4*)pos, order, n, ctx->d_spos, s);
  if (e == hipSuccess) e = hipEventRecord(ev[1], s);
  if (e == hipSuccess)
    e = launch_deposit_ex(P, sorted ? ctx->d_spos : (const float4*)pos, n, ctx->d_rot, ctx->d_off,
                          ctx->d_frot, (unsigned long long*)dep, cache ? ctx->d_cells : nullptr,
                          sorted, ctx->tune, s, sorted ? nullptr : ensure_cpos(ctx, n));
  if (e == hipSuccess) e = hipEventRecord(ev[2], s);
  if (e == hipSuccess) e = solve_dispatch(ctx, dep, ctx->d_force, 1, s);
  if (e == hipSuccess) e = hipEventRecord(ev[3], s);
  if (e == hipSuccess) {
    if (cache) {
      e = launch_gather_cached(P, (float4*)pos, (float4*)vel, n, order, ctx->d_cells, ctx->d_force,
                               s);
      if (e == hipSuccess) e = launch_fill_zero(old, cells, s);
    } else {
      e = launch_gather_update(P, (float4*)pos, (float4*)vel, n, ctx->d_rot, ctx->d_off,
                               ctx->d_frot, ctx->d_force, 1, nullptr, 0, order, s, old, cells,
                               ps, vs);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(ev[4], s);
  if (e == hipSuccess && state) e = launch_state_export(vs, order, n, (float4*)vel, s);
  if (e == hipSuccess) swap_grids(ctx);
  if (e == hipSuccess) e = hipEventSynchronize(ev[4]);
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventElapsedTime(&phase_ms[i], ev[i], ev[i + 1]);
  for (int i = 0; i < 5; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (e != hipSuccess) ctx->spare_dirty = true;
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_step_profiled");
}

int mog_deposit(mog_ctx* ctx, const float* pos, int64_t n, int64_t* grids_out, void* stream) {
  int rc = check_particles(ctx, pos, nullptr, n, false);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (!grids_out) return fail(ctx, MOG_EINVAL, "grids_out is NULL");
  if (((uintptr_t)grids_out & 7u) != 0) return fail(ctx, MOG_EINVAL, "grids_out must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const long long cells = (long long)ctx->P.k * ctx->P.cells;
  float4* const cpos = ensure_cpos(ctx, n);
  hipError_t e = launch_fill_zero((long long*)grids_out, cells, s);
  if (e == hipSuccess)
    e = launch_deposit(ctx->P, (const float4*)pos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                       (unsigned long long*)grids_out, ctx->tune, s, cpos);
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_deposit");
}

int mog_solve(mog_ctx* ctx, const int64_t* grids, float* force_out, void* stream) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (!grids || !force_out) return fail(ctx, MOG_EINVAL, "grids/force_out is NULL");
  if (((uintptr_t)grids & 7u) != 0) return fail(ctx, MOG_EINVAL, "grids must be 8-byte aligned");
  if (!aligned16(force_out)) return fail(ctx, MOG_EINVAL, "force_out must be 16-byte aligned");
  hipError_t e = solve_dispatch(ctx, (const long long*)grids, (float4*)force_out, 0,
                                (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_solve");
}

int mog_gather(mog_ctx* ctx, const float* pos, int64_t n, const float* force, float* avg_out,
               void* stream) {
  int rc = check_particles(ctx, pos, nullptr, n, false);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (n == 0) return MOG_OK;
  if (!force || !avg_out) return fail(ctx, MOG_EINVAL, "force/avg_out is NULL");
  if (!aligned16(force) || !aligned16(avg_out))
    return fail(ctx, MOG_EINVAL, "force/avg_out must be 16-byte aligned");
  hipError_t e = launch_gather_update(ctx->P, (float4*)pos, nullptr, n, ctx->d_rot, ctx->d_off,
                                      ctx->d_frot, (const float4*)force, 0, (float4*)avg_out, 1,
                                      nullptr, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_gather");
}

int mog_update(mog_ctx* ctx, float* pos, float* vel, const float* avg, int64_t n, void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (n == 0) return MOG_OK;
  if (!avg) return fail(ctx, MOG_EINVAL, "avg_force is NULL");
  if (!aligned16(avg)) return fail(ctx, MOG_EINVAL, "avg_force must be 16-byte aligned");
  hipError_t e = launch_update(ctx->P, (float4*)pos, (float4*)vel, (const float4*)avg, n,
                               (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_update");
}

int mog_set_integrator(mog_ctx* ctx, double dt, double damping, double timescale) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (!isfinite(dt) || !isfinite(damping) || !isfinite(timescale))
    return fail(ctx, MOG_EINVAL, "dt, damping and timescale must be finite");
  ctx->cfg.dt = dt;
  ctx->cfg.damping = damping;
  ctx->cfg.timescale = timescale;
  ctx->P.dtp = (float)(dt * timescale);
  ctx->P.damping = (float)damping;
  drop_graphs(ctx);  // kernel arguments of a captured step changed
  return MOG_OK;
}

int mog_init_perlin(float* pos, float* vel, int64_t n, double world_size, double noise_scale,
                    double density_contrast, uint32_t seed, double base_mass, double spawn_buffer,
                    int32_t spawn_shape, void* stream) {
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (n == 0) return MOG_OK;
  if (!pos || !vel) return fail(nullptr, MOG_EINVAL, "pos/vel is NULL");
  if (!aligned16(pos) || !aligned16(vel)) return fail(nullptr, MOG_EINVAL, "pos/vel must be 16-byte aligned");
  if (!(isfinite(world_size) && world_size > 0.0) || !isfinite(noise_scale) ||
      !isfinite(density_contrast) || !isfinite(base_mass) || !isfinite(spawn_buffer))
    return fail(nullptr, MOG_EINVAL, "parameters must be finite (world_size > 0)");
  if (spawn_shape != 0 && spawn_shape != 1) return fail(nullptr, MOG_EINVAL, "spawn_shape must be 0 or 1");
  hipError_t e = launch_init_perlin((float4*)pos, (float4*)vel, n, (float)world_size, (float)noise_scale,
                                    (float)density_contrast, seed, (float)base_mass,
                                    (float)spawn_buffer, spawn_shape, (hipStream_t)stream);
  if (e != hipSuccess) return fail(nullptr, MOG_EHIP, "mog_init_perlin: %s", hipGetErrorString(e));
  return MOG_OK;
}

int mog_closure_features(mog_ctx* ctx, const float* pos, int64_t n, double rho0, float* feat,
                         float* mean, void* stream) {
  int rc = check_particles(ctx, pos, nullptr, n, false);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (n == 0) return MOG_OK;
  if (!feat || !mean || !aligned16(feat) || !aligned16(mean))
    return fail(ctx, MOG_EINVAL, "feat/mean must be 16-byte aligned device pointers");
  if (!(isfinite(rho0) && rho0 > 0.0)) return fail(ctx, MOG_EINVAL, "rho0 must be finite and > 0");
  hipStream_t s = (hipStream_t)stream;
  const Params& P = ctx->P;
  if (!ctx->d_gtab_dc) {
    // the spectral Green's table divided by the NGP window
    // sinc(k_x h/2) sinc(k_y h/2) sinc(k_z h/2) (Nyquist-positive k)
    const int g = P.g;
    std::vector<float> gt((size_t)g * g * g), dc((size_t)g * g * g);
    std::vector<float2> tw(g);
    std::vector<float> kd(g);
    build_tables(&ctx->cfg, tw.data(), gt.data(), kd.data());
    auto win = [&](int i) {
      const int sh = i <= g / 2 ? i : i - g;
      const double x = 0.5 * (double)sh * 2.0 * M_PI / ctx->cfg.world_size * (double)P.h;
      return sh == 0 ? 1.0 : sin(x) / x;
    };
    for (int iz = 0; iz < g; ++iz)
      for (int iy = 0; iy < g; ++iy)
        for (int ix = 0; ix < g; ++ix) {
          const size_t at = ((size_t)iz * g + iy) * g + ix;
          dc[at] = (float)((double)gt[at] / (win(ix) * win(iy) * win(iz)));
        }
    if (dalloc(&ctx->d_gtab_dc, dc.size()) != hipSuccess ||
        dalloc(&ctx->d_force_dc, (size_t)P.k * P.cells) != hipSuccess ||
        dalloc(&ctx->d_fm, (size_t)P.k * P.cells) != hipSuccess ||
        dalloc(&ctx->d_blk, (size_t)P.k * P.cells * 8) != hipSuccess) {
      (void)hipFree(ctx->d_gtab_dc);
      (void)hipFree(ctx->d_force_dc);
      (void)hipFree(ctx->d_fm);
      (void)hipFree(ctx->d_blk);
      ctx->d_gtab_dc = nullptr;
      ctx->d_force_dc = nullptr;
      ctx->d_fm = nullptr;
      ctx->d_blk = nullptr;
      return fail(ctx, MOG_ENOMEM, "mog_closure_features: deconvolved grids");
    }
    hipError_t ec = hipMemcpy(ctx->d_gtab_dc, dc.data(), dc.size() * sizeof(float),
                              hipMemcpyHostToDevice);
    if (ec != hipSuccess) return hip_fail(ctx, ec, "mog_closure_features");
  }
  hipError_t e = launch_fill_zero(ctx->d_grids, (long long)P.k * P.cells, s);
  // the few grids of a closure context spread the workgroup target over
  // many small chunks: keep each workgroup's pairs well above its counters,
  // whose flush otherwise costs as much as the deposit
  DepositTune dt = ctx->tune;
  dt.min_chunk = 4ll * P.cells;
  if (e == hipSuccess)
    e = launch_deposit(P, (const float4*)pos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                       (unsigned long long*)ctx->d_grids, dt, s);
  // the plain and the deconvolved force grids: one launch of both solves
  // where the grid solve of this size runs them (g = 16, vector FFT path)
  const bool pair = P.g == 16 && !ctx->spectral_mfma && ctx->cfg.precision == MOG_PRECISION_FP32;
  if (e == hipSuccess && pair)
    e = launch_solve_pair(P, ctx->d_grids, ctx->d_force, ctx->d_gtab, ctx->d_conv,
                          ctx->d_force_dc, ctx->d_gtab_dc, nullptr, ctx->d_kd, ctx->need_x,
                          ctx->d_rot, 1, s);
  if (e == hipSuccess && !pair) e = solve_dispatch(ctx, ctx->d_grids, ctx->d_force, 1, s);
  if (e == hipSuccess && !pair)
    e = launch_solve(P, ctx->d_grids, ctx->d_S1, ctx->s1_elems, ctx->d_S2, ctx->d_force_dc,
                     ctx->d_tw,
                     ctx->d_gtab_dc, ctx->d_kd, nullptr, ctx->need_x, ctx->d_rot, 1, s);
  const int* order = nullptr;
  if (e == hipSuccess) e = refresh_order(ctx, pos, n, s, &order);
  if (e == hipSuccess)
    e = launch_closure_features(P, (const float4*)pos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                                ctx->d_force, ctx->d_force_dc, ctx->d_grids, (float)rho0, order,
                                ctx->d_fm, ctx->d_blk, (float4*)feat, (float4*)mean, s);
  return e == hipSuccess ? MOG_OK : hip_fail(ctx, e, "mog_closure_features");
}

int mog_diagnostics(mog_ctx* ctx, const float* pos, const float* vel, int64_t n, double* out8,
                    void* stream) {
  int rc = check_particles(ctx, pos, vel, n, true);
  if (rc != MOG_OK) return rc;
  if (!ctx->have_transforms) return fail(ctx, MOG_ENOTRANSFORMS, "mog_set_transforms not called");
  if (!out8 || ((uintptr_t)out8 & 7u) != 0)
    return fail(ctx, MOG_EINVAL, "out8 must be an 8-byte aligned device pointer");
  hipStream_t s = (hipStream_t)stream;
  const Params& P = ctx->P;
  const int g = P.g;
  if (!ctx->d_diag_tab) {
    // FP64 twiddles and T(k) = G_hat(k) / (2 g^3 10^6): the Green's
    // function of build_tables (same float32 k2 mask) in milli-mass units
    std::vector<double2> tw(g);
    for (int j = 0; j < g; ++j) {
      const double th = 2.0 * M_PI * (double)j / (double)g;
      tw[j] = make_double2(cos(th), -sin(th));
    }
    const double W = ctx->cfg.world_size;
    auto sh = [g](int i) { return i <= g / 2 ? i : i - g; };
    const double scale = 1.0 / (2.0 * (double)g * g * g * 1e6);
    std::vector<double> tab((size_t)g * g * g);
    for (int iz = 0; iz < g; ++iz)
      for (int iy = 0; iy < g; ++iy)
        for (int ix = 0; ix < g; ++ix) {
          double v = 0.0;
          if ((ix | iy | iz) != 0 && green_mask32(ix, iy, iz, g, W)) {
            const double kx = sh(ix) * 2.0 * M_PI / W, ky = sh(iy) * 2.0 * M_PI / W,
                         kz = sh(iz) * 2.0 * M_PI / W;
            v = -2.0 * M_PI * ctx->cfg.G / (kx * kx + ky * ky + kz * kz) * scale;
          }
          tab[((size_t)iz * g + iy) * g + ix] = v;
        }
    if (dalloc(&ctx->d_diag_tw, (size_t)g) != hipSuccess ||
        dalloc(&ctx->d_diag_tab, tab.size()) != hipSuccess ||
        dalloc(&ctx->d_diag_scratch, diag_scratch_doubles(P.k, g)) != hipSuccess) {
      (void)hipFree(ctx->d_diag_tw);
      (void)hipFree(ctx->d_diag_tab);
      (void)hipFree(ctx->d_diag_scratch);
      ctx->d_diag_tw = nullptr;
      ctx->d_diag_tab = nullptr;
      ctx->d_diag_scratch = nullptr;
      return fail(ctx, MOG_ENOMEM, "mog_diagnostics: tables");
    }
    hipError_t ec = hipMemcpy(ctx->d_diag_tw, tw.data(), tw.size() * sizeof(double2),
                              hipMemcpyHostToDevice);
    if (ec == hipSuccess)
      ec = hipMemcpy(ctx->d_diag_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    if (ec != hipSuccess) {
      (void)hipFree(ctx->d_diag_tab);
      ctx->d_diag_tab = nullptr;
      return hip_fail(ctx, ec, "mog_diagnostics");
    }
  }
  // the density of the current positions in the spare grids (all zero
  // between steps), which are cleared again afterwards
  const long long cells = (long long)P.k * P.cells;
  hipError_t e = clean_spare(ctx, s);
  if (e == hipSuccess)
    e = launch_deposit(P, (const float4*)pos, n, ctx->d_rot, ctx->d_off, ctx->d_frot,
                       (unsigned long long*)ctx->d_grids_next, ctx->tune, s);
  if (e == hipSuccess)
    e = launch_diagnostics((const float4*)pos, (const float4*)vel, n, ctx->d_grids_next, P.k, g,
                           ctx->d_diag_tw, ctx->d_diag_tab, ctx->d_diag_scratch, out8, s);
  if (e == hipSuccess) e = launch_fill_zero(ctx->d_grids_next, cells, s);
  if (e != hipSuccess) {
    ctx->spare_dirty = true;
    return hip_fail(ctx, e, "mog_diagnostics");
  }
  return MOG_OK;
}

int mog_mlp_param_count(void) { return (int)mlp_param_count(); }

int mog_mlp_forward(const float* params, const float* feat, const float* mean, int64_t n,
                    float* force, int32_t bf16, void* stream) {
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (n == 0) return MOG_OK;
  if (!params || !feat || !mean || !force) return fail(nullptr, MOG_EINVAL, "NULL pointer");
  if (!aligned16(feat) || !aligned16(mean) || !aligned16(force))
    return fail(nullptr, MOG_EINVAL, "feat/mean/force must be 16-byte aligned");
  hipError_t e = launch_mlp_forward(params, (const float4*)feat, (const float4*)mean, n,
                                    (float4*)force, bf16 != 0, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_mlp_forward: %s", hipGetErrorString(e));
}

int mog_mlp_train_grad(const float* params, const float* feat, const float* mean,
                       const float* target, int64_t n, double wnorm, float* grad, float* loss,
                       int32_t bf16, void* stream) {
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (n == 0) return MOG_OK;
  if (!params || !feat || !mean || !target || !grad || !loss)
    return fail(nullptr, MOG_EINVAL, "NULL pointer");
  if (!aligned16(feat) || !aligned16(mean) || !aligned16(target))
    return fail(nullptr, MOG_EINVAL, "feat/mean/target must be 16-byte aligned");
  if (!(isfinite(wnorm) && wnorm > 0.0)) return fail(nullptr, MOG_EINVAL, "wnorm must be > 0");
  hipError_t e = launch_mlp_train(params, (const float4*)feat, (const float4*)mean,
                                  (const float4*)target, n, (float)wnorm, grad, loss, bf16 != 0,
                                  (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_mlp_train_grad: %s", hipGetErrorString(e));
}

int mog_direct_sum(const float* pos, int64_t n, const int32_t* targets, int32_t nt,
                   double world_size, double G_eff, double softening, double* out, void* stream) {
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (nt < 0) return fail(nullptr, MOG_EINVAL, "nt must be >= 0");
  if (nt == 0) return MOG_OK;
  if (!pos || !targets || !out) return fail(nullptr, MOG_EINVAL, "NULL pointer");
  if (!aligned16(pos)) return fail(nullptr, MOG_EINVAL, "pos must be 16-byte aligned");
  if (!(isfinite(world_size) && world_size > 0.0) || !isfinite(G_eff) || !isfinite(softening))
    return fail(nullptr, MOG_EINVAL, "world_size > 0, G_eff and softening finite");
  hipError_t e = launch_direct_sum((const float4*)pos, n, targets, nt, world_size, G_eff,
                                   softening * softening, out, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_direct_sum: %s", hipGetErrorString(e));
}

int mog_mlp_adam(float* params, const float* grad, float* m1, float* m2, double lr, double beta1,
                 double beta2, double eps, int32_t step, void* stream) {
  if (!params || !grad || !m1 || !m2) return fail(nullptr, MOG_EINVAL, "NULL pointer");
  if (step < 1) return fail(nullptr, MOG_EINVAL, "step must be >= 1");
  hipError_t e = launch_adam(params, grad, m1, m2, (float)lr, (float)beta1, (float)beta2, (float)eps,
                             step, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_mlp_adam: %s", hipGetErrorString(e));
}

int mog_get_tuning(mog_ctx* ctx, mog_tuning* out) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (!out) return fail(ctx, MOG_EINVAL, "out is NULL");
  memset(out, 0, sizeof *out);
  out->deposit_threads = ctx->tune.threads;
  out->deposit_grids_per_wg = ctx->tune.grids_per_wg;
  out->deposit_wg_target = ctx->tune.wg_target;
  out->resort_every = ctx->resort_every;
  out->sorted_deposit = ctx->sorted_deposit ? 1 : 0;
  out->cell_cache = ctx->cell_cache ? 1 : 0;
  out->spectral_mfma = ctx->spectral_mfma ? 1 : 0;
  out->deposit_prepass = ctx->tune.prepass;
  return MOG_OK;
}

int mog_set_tuning(mog_ctx* ctx, const mog_tuning* t) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (!t) return fail(ctx, MOG_EINVAL, "tuning is NULL");
  if (t->deposit_threads < 64 || t->deposit_threads > 1024 || (t->deposit_threads % 64) != 0)
    return fail(ctx, MOG_EINVAL, "deposit_threads must be a multiple of 64 in [64, 1024]");
  const int G = t->deposit_grids_per_wg;
  if (G != 1 && G != 2 && G != 4 && G != 8)
    return fail(ctx, MOG_EINVAL, "deposit_grids_per_wg must be 1, 2, 4 or 8");
  if (t->deposit_wg_target < 1) return fail(ctx, MOG_EINVAL, "deposit_wg_target must be >= 1");
  if (t->resort_every < 0) return fail(ctx, MOG_EINVAL, "resort_every must be >= 0");
  if ((t->sorted_deposit | t->cell_cache | t->spectral_mfma) & ~1)
    return fail(ctx, MOG_EINVAL, "flags must be 0 or 1");
  if (t->deposit_prepass < -1 || t->deposit_prepass > 1)
    return fail(ctx, MOG_EINVAL, "deposit_prepass must be -1, 0 or 1");
  if (t->cell_cache && !t->sorted_deposit)
    return fail(ctx, MOG_EINVAL, "cell_cache needs sorted_deposit");
  ctx->tune.threads = t->deposit_threads;
  ctx->tune.grids_per_wg = G;
  ctx->tune.wg_target = t->deposit_wg_target;
  ctx->tune.prepass = t->deposit_prepass;
  const bool realloc = (t->sorted_deposit != 0) != ctx->sorted_deposit;
  ctx->resort_every = t->resort_every;
  ctx->sorted_deposit = t->sorted_deposit != 0;
  ctx->cell_cache = t->cell_cache != 0;
  ctx->spectral_mfma = t->spectral_mfma != 0;
  if (realloc) free_order(ctx);  // buffers follow the sorted-deposit choice
  ctx->order_pos = nullptr;      // rebuild the order on the next step
  drop_graphs(ctx);
  return MOG_OK;
}

const int64_t* mog_grids_ptr(mog_ctx* ctx) { return ctx ? (const int64_t*)ctx->d_grids : nullptr; }

const float* mog_force_ptr(mog_ctx* ctx) { return ctx ? (const float*)ctx->d_force : nullptr; }

int mog_batch_create(const mog_config* cfgs, const int64_t* n, int32_t n_sims, mog_batch** out) {
  if (!out) return fail(nullptr, MOG_EINVAL, "out is NULL");
  *out = nullptr;
  if (!cfgs || !n) return fail(nullptr, MOG_EINVAL, "cfgs/n is NULL");
  if (n_sims < 1 || n_sims > 65536) return fail(nullptr, MOG_EINVAL, "n_sims must be in [1, 65536]");
  mog_batch* b = new mog_batch();
  b->nsims = n_sims;
  b->cfg.assign(cfgs, cfgs + n_sims);
  b->sims.resize(n_sims);
  b->have_transforms.assign(n_sims, 0);
  long long p_off = 0, cell_off = 0, conv_off = 0, blk_off = 0, cls_slabs[kBatchClasses] = {};
  int grid_off = 0, tab_off = 0, line_off = 0;
  std::vector<int> cls_sim[kBatchClasses];
  std::vector<long long> cls_off[kBatchClasses];
  int cls_gmax[kBatchClasses] = {};
  for (int i = 0; i < n_sims; ++i) {
    const mog_config* c = &cfgs[i];
    int rc = validate(c);
    if (rc != MOG_OK) {
      delete b;
      return fail(nullptr, rc, "sim %d: %s", i, g_create_error.c_str());
    }
    if (c->grid_size > 32 || c->precision != MOG_PRECISION_FP32) {
      delete b;
      return fail(nullptr, MOG_EUNSUPPORTED, "sim %d: the batch runs grid_size <= 32 in fp32", i);
    }
    if (n[i] < 0) {
      delete b;
      return fail(nullptr, MOG_EINVAL, "sim %d: n must be >= 0", i);
    }
    BatchSim& sm = b->sims[i];
    sm.P = make_params(c);
    sm.p_off = p_off;
    sm.n = n[i];
    const int ci = ilog2(c->grid_size) - 1;
    sm.cell_off = cell_off;
    sm.slab_off = cls_slabs[ci];
    cls_sim[ci].push_back(i);
    cls_off[ci].push_back(cls_slabs[ci]);
    cls_gmax[ci] = std::max(cls_gmax[ci], c->grid_size);
    sm.grid_off = grid_off;
    sm.tab_off = tab_off;
    sm.line_off = line_off;
    sm.need_x = !(c->compat == MOG_COMPAT_REFERENCE && aliases_x_with_z(c->grid_size));
    const bool conv = c->solver == MOG_SOLVER_CONVOLUTION && c->compat == MOG_COMPAT_PHYSICAL;
    sm.conv_off = conv ? conv_off : -1;
    sm.blk_off = blk_off;
    blk_off += (n[i] + kBatchGatherBlock - 1) / kBatchGatherBlock;
    const long long g = c->grid_size, cells = (long long)c->n_grids * g * g * g;
    p_off += n[i];
    cell_off += cells;
    cls_slabs[ci] += (long long)c->n_grids * g;
    grid_off += c->n_grids;
    tab_off += (int)(g * g * g);
    line_off += (int)g;
    if (conv) conv_off += conv_table_elems((int)g);
  }
  if (p_off > 2147483647LL || cell_off > 2147483647LL) {
    delete b;
    return fail(nullptr, MOG_EINVAL, "total particles and cells must stay below 2^31");
  }
  b->ntotal = p_off;
  b->cells_total = cell_off;
  b->gather_blocks = blk_off;
  std::vector<int> all_sim;
  std::vector<long long> all_off;
  for (int ci = 0; ci < kBatchClasses; ++ci) {
    b->cls[ci].start = (int)all_sim.size();
    b->cls[ci].count = (int)cls_sim[ci].size();
    b->cls[ci].gmax = cls_gmax[ci];
    b->cls[ci].slabs = cls_slabs[ci];
    all_sim.insert(all_sim.end(), cls_sim[ci].begin(), cls_sim[ci].end());
    all_off.insert(all_off.end(), cls_off[ci].begin(), cls_off[ci].end());
  }
  b->grids_total = grid_off;
  b->kmax = 0;
  for (int i = 0; i < n_sims; ++i) b->kmax = std::max(b->kmax, cfgs[i].n_grids);
  // solve windows of the classes g <= 8 (the slab passes): whole grids of
  // one class, at most kBatchWindowCells cells each, laid out in the
  // window's scratch in slab order.  The g = 16 and g = 32 classes run one
  // launch each over all of their grids (one workgroup per grid).
  std::vector<long long> win_soff;
  long long win_max = 0;
  for (int ci = 0; ci < kBatchClasses; ++ci) {
    if (batch_class_fused(cls_gmax[ci])) continue;
    long long used = 0;
    for (size_t jj = 0; jj < cls_sim[ci].size(); ++jj) {
      const mog_config* c = &cfgs[cls_sim[ci][jj]];
      const long long g = c->grid_size, g3 = g * g * g;
      long long placed_in = -1;  // window that holds this simulation's offset
      for (int grid = 0; grid < c->n_grids; ++grid) {
        const long long slab = cls_off[ci][jj] + (long long)grid * g;
        if (b->windows.empty() || b->windows.back().cls != ci || used + g3 > kBatchWindowCells) {
          b->windows.push_back({ci, slab, 0});
          win_soff.resize(b->windows.size() * (size_t)n_sims, 0);
          used = 0;
        }
        const long long w = (long long)b->windows.size() - 1;
        if (placed_in != w) {  // the simulation's first grid in this window
          win_soff[(size_t)w * n_sims + jj] = used - (long long)grid * g3;
          placed_in = w;
        }
        b->windows.back().slabs += g;
        used += g3;
        win_max = std::max(win_max, used);
      }
    }
  }
  b->deposit_chunk = 131072;
  b->deposit_threads = 1024;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(dalloc(&b->d_sims, n_sims));
  A(dalloc(&b->d_rot, (size_t)grid_off * 12));
  A(dalloc(&b->d_off, (size_t)grid_off * 4));
  A(dalloc(&b->d_frot, (size_t)grid_off * 12));
  A(dalloc(&b->d_grids, (size_t)cell_off));
  const long long g32_grids = cls_gmax[4] == 32 ? cls_slabs[4] / 32 : 0;
  b->s1_elems = std::max(win_max, g32_grids > 0 ? rh_scratch_elems(g32_grids) : 0LL);
  A(dalloc(&b->d_S1, (size_t)std::max(b->s1_elems, 1LL)));
  A(dalloc(&b->d_S2, (size_t)std::max(win_max, 1LL)));
  A(dalloc(&b->d_win_soff, win_soff.size() > 0 ? win_soff.size() : 1));
  A(dalloc(&b->d_force, (size_t)cell_off));
  A(dalloc(&b->d_tw, (size_t)line_off));
  A(dalloc(&b->d_gtab, (size_t)tab_off));
  A(dalloc(&b->d_kd, (size_t)line_off));
  A(dalloc(&b->d_conv, (size_t)(conv_off > 0 ? conv_off : 1)));
  A(dalloc(&b->d_counter, 1));
  A(dalloc(&b->d_cpos, (size_t)std::max(p_off, 1LL)));
  A(dalloc(&b->d_cls_sim, (size_t)n_sims));
  A(dalloc(&b->d_cls_off, (size_t)n_sims));
  b->resort_every = 16;
  size_t tmp = 0;
  A(batch_order_temp_bytes(p_off, n_sims, &tmp));
  b->sort_tmp_bytes = tmp;
  A(dalloc(&b->d_order, (size_t)p_off));
  A(dalloc(&b->d_idx, (size_t)p_off));
  A(dalloc(&b->d_keys_a, (size_t)p_off));
  A(dalloc(&b->d_keys_b, (size_t)p_off));
  A(hipMalloc(&b->d_sort_tmp, tmp > 0 ? tmp : 16));
  if (e != hipSuccess) {
    batch_release(b);
    delete b;
    return fail(nullptr, MOG_ENOMEM, "device allocation failed for the batch (%lld cells)", cell_off);
  }
  std::vector<float2> tw(line_off), conv(conv_off > 0 ? conv_off : 1);
  std::vector<float> gtab(tab_off), kd(line_off);
  for (int i = 0; i < n_sims; ++i) {
    const BatchSim& sm = b->sims[i];
    build_tables(&cfgs[i], tw.data() + sm.line_off, gtab.data() + sm.tab_off, kd.data() + sm.line_off);
    if (sm.conv_off >= 0) build_conv_tables(&cfgs[i], conv.data() + sm.conv_off);
  }
  A(hipMemcpy(b->d_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_gtab, gtab.data(), gtab.size() * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_kd, kd.data(), kd.size() * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_conv, conv.data(), conv.size() * sizeof(float2), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_sims, b->sims.data(), b->sims.size() * sizeof(BatchSim), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_cls_sim, all_sim.data(), all_sim.size() * sizeof(int), hipMemcpyHostToDevice));
  A(hipMemcpy(b->d_cls_off, all_off.data(), all_off.size() * sizeof(long long),
              hipMemcpyHostToDevice));
  if (!win_soff.empty())
    A(hipMemcpy(b->d_win_soff, win_soff.data(), win_soff.size() * sizeof(long long),
                hipMemcpyHostToDevice));
  A(batch_build_tasks(b));
  if (e != hipSuccess) {
    batch_release(b);
    delete b;
    return fail(nullptr, MOG_EHIP, "batch tables: %s", hipGetErrorString(e));
  }
  *out = b;
  return MOG_OK;
}

void mog_batch_destroy(mog_batch* b) {
  if (!b) return;
  batch_release(b);
  delete b;
}

int mog_batch_set_transforms(mog_batch* b, int32_t sim, const float* rot_k12, const float* off_k4) {
  if (!b) return fail(nullptr, MOG_EINVAL, "batch is NULL");
  if (sim < 0 || sim >= b->nsims) return fail(nullptr, MOG_EINVAL, "sim %d out of range", sim);
  if (!rot_k12 || !off_k4) return fail(nullptr, MOG_EINVAL, "rot/off is NULL");
  BatchSim& sm = b->sims[sim];
  const int k = sm.P.k;
  double rmax = 0.0, omax = 0.0;
  bool finite = true;
  for (int i = 0; i < k * 12; ++i) {
    finite = finite && isfinite(rot_k12[i]);
    rmax = fmax(rmax, fabs((double)rot_k12[i]));
  }
  for (int i = 0; i < k; ++i)
    for (int m = 0; m < 3; ++m) {
      finite = finite && isfinite(off_k4[4 * i + m]);
      omax = fmax(omax, fabs((double)off_k4[4 * i + m]));
    }
  if (!finite) return fail(nullptr, MOG_EINVAL, "rotations/offsets must be finite");
  std::vector<float> frot((size_t)k * 12);
  for (int i = 0; i < k; ++i)
    make_fast_row(rot_k12 + 12 * i, off_k4 + 4 * i, sm.P.half, sm.P.h, frot.data() + 12 * i);
  const double margin = fast_margin(b->cfg[sim].world_size, sm.P.g, rmax, omax, sm.P.CB);
  sm.P.lim = margin < 0.05 ? (float)(0.5 - margin) : -1.0f;
  sm.P.shift_ok = shift_flags(b->cfg[sim].world_size, sm.P.g, sm.P.lg, rmax, omax, sm.P.CB);
  const size_t g0 = (size_t)sm.grid_off;
  hipError_t e = hipMemcpy(b->d_rot + 12 * g0, rot_k12, (size_t)k * 12 * sizeof(float),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_off + 4 * g0, off_k4, (size_t)k * 4 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_frot + 12 * g0, frot.data(), frot.size() * sizeof(float),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_sims + sim, &sm, sizeof(BatchSim), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(nullptr, MOG_EHIP, "mog_batch_set_transforms: %s", hipGetErrorString(e));
  b->have_transforms[sim] = 1;
  return MOG_OK;
}

int mog_batch_step(mog_batch* b, float* pos, float* vel, void* stream) {
  int rc = batch_check(b, pos, vel);
  if (rc != MOG_OK) return rc;
  if (b->ntotal == 0) return MOG_OK;
  hipError_t e = batch_refresh_order(b, pos, (hipStream_t)stream);
  if (e == hipSuccess) e = batch_enqueue(b, pos, vel, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_batch_step: %s", hipGetErrorString(e));
}

int mog_batch_run(mog_batch* b, float* pos, float* vel, int32_t nsteps, int32_t use_graph,
                  void* stream) {
  int rc = batch_check(b, pos, vel);
  if (rc != MOG_OK) return rc;
  if (nsteps < 0) return fail(nullptr, MOG_EINVAL, "nsteps must be >= 0");
  if (b->ntotal == 0 || nsteps == 0) return MOG_OK;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  // the state in visiting order, as in mog_run: imported at the start and
  // after each rebuild of the order, velocities written back at each
  // rebuild and at the end
  const bool state = nsteps >= kStateMinSteps && batch_order(b) && batch_ensure_state(b);
  float4* const ps = state ? b->d_ps : nullptr;
  float4* const vs = state ? b->d_vs : nullptr;
  bool imported = false;
  auto prepare = [&](int i) -> hipError_t {
    hipError_t ee = hipSuccess;
    const bool rebuild = b->order_pos != pos || b->steps_since_order >= b->resort_every;
    if (state && imported && rebuild) {
      if ((ee = launch_state_export(vs, b->d_order, b->ntotal, (float4*)vel, s)) != hipSuccess)
        return ee;
      imported = false;
    }
    if ((ee = batch_refresh_order(b, pos, s)) != hipSuccess) return ee;
    if (state && (i == 0 || rebuild)) {
      if ((ee = launch_state_import((const float4*)pos, (const float4*)vel, b->d_order, b->ntotal,
                                    ps, vs, s)) != hipSuccess)
        return ee;
      imported = true;
    }
    return hipSuccess;
  };
  auto finish = [&](hipError_t ee, const char* what) -> int {
    if (state && imported) {
      const hipError_t ex = launch_state_export(vs, b->d_order, b->ntotal, (float4*)vel, s);
      if (ee == hipSuccess) ee = ex;
    }
    if (ee != hipSuccess) return fail(nullptr, MOG_EHIP, "%s: %s", what, hipGetErrorString(ee));
    return MOG_OK;
  };
  if (!use_graph) {
    for (int i = 0; i < nsteps; ++i)
      if ((e = prepare(i)) != hipSuccess || (e = batch_enqueue(b, pos, vel, s, ps, vs)) != hipSuccess)
        return finish(e, "mog_batch_run");
    return finish(hipSuccess, "mog_batch_run");
  }
  if (!b->gexec || b->g_pos != pos || b->g_vel != vel || b->g_ps != ps) {
    batch_drop_graph(b);
    if (!b->cap_stream &&
        (e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess)
      return fail(nullptr, MOG_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    if ((e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return fail(nullptr, MOG_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    const hipError_t ee = batch_enqueue(b, pos, vel, b->cap_stream, ps, vs);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &graph);
    if (ee != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(nullptr, MOG_EHIP, "capture: %s", hipGetErrorString(ee));
    }
    if (e != hipSuccess) return fail(nullptr, MOG_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      b->gexec = nullptr;
      return fail(nullptr, MOG_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->g_pos = pos;
    b->g_vel = vel;
    b->g_ps = ps;
  }
  for (int i = 0; i < nsteps; ++i) {
    if ((e = prepare(i)) != hipSuccess) return finish(e, "mog_batch_run");
    if ((e = hipGraphLaunch(b->gexec, s)) != hipSuccess) return finish(e, "hipGraphLaunch");
  }
  return finish(hipSuccess, "mog_batch_run");
}

int mog_batch_set_tuning(mog_batch* b, int64_t deposit_chunk, int32_t deposit_threads) {
  if (!b) return fail(nullptr, MOG_EINVAL, "batch is NULL");
  if (deposit_chunk < 256 || deposit_chunk > (1LL << 24))
    return fail(nullptr, MOG_EINVAL, "deposit_chunk must be in [256, 2^24]");
  if (deposit_threads != 256 && deposit_threads != 512 && deposit_threads != 1024)
    return fail(nullptr, MOG_EINVAL, "deposit_threads must be 256, 512 or 1024");
  b->deposit_chunk = deposit_chunk;
  b->deposit_threads = deposit_threads;
  batch_drop_graph(b);
  hipError_t e = batch_build_tasks(b);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog_batch_set_tuning: %s", hipGetErrorString(e));
}

int mog_batch_get_tuning(const mog_batch* b, int64_t* deposit_chunk, int32_t* deposit_threads,
                         int32_t* n_tasks) {
  if (!b) return fail(nullptr, MOG_EINVAL, "batch is NULL");
  if (deposit_chunk) *deposit_chunk = b->deposit_chunk;
  if (deposit_threads) *deposit_threads = b->deposit_threads;
  if (n_tasks) *n_tasks = b->ntasks;
  return MOG_OK;
}

int mog2d_create(int32_t n_grids, int32_t grid_size, double world_size, double G,
                 const float* rot_k12, const float* off_k4, mog2d_ctx** out) {
  if (!out) return fail(nullptr, MOG_EINVAL, "out is NULL");
  *out = nullptr;
  if (n_grids < 1 || n_grids > 65535) return fail(nullptr, MOG_EINVAL, "n_grids must be in [1, 65535]");
  if (!is_pow2(grid_size) || grid_size < 2 || grid_size > 64)
    return fail(nullptr, MOG_EINVAL, "grid_size must be a power of two in [2, 64]");
  if (!(isfinite(world_size) && world_size > 0.0) || !isfinite(G))
    return fail(nullptr, MOG_EINVAL, "world_size must be finite and > 0, G finite");
  if (!rot_k12 || !off_k4) return fail(nullptr, MOG_EINVAL, "rot/off is NULL");
  for (int i = 0; i < n_grids * 12; ++i)
    if (!isfinite(rot_k12[i])) return fail(nullptr, MOG_EINVAL, "rotations must be finite");
  for (int i = 0; i < n_grids * 4; ++i)
    if (!isfinite(off_k4[i])) return fail(nullptr, MOG_EINVAL, "offsets must be finite");
  mog2d_ctx* c = new mog2d_ctx();
  c->k = n_grids;
  c->g = grid_size;
  c->lg = ilog2(grid_size);
  c->W = world_size;
  c->G = G;
  const size_t g = (size_t)grid_size, cells = (size_t)n_grids * g * g;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(dalloc(&c->d_rot, (size_t)n_grids * 12));
  A(dalloc(&c->d_off, (size_t)n_grids * 4));
  A(dalloc(&c->d_frot, (size_t)n_grids * 12));
  A(dalloc(&c->d_tw, g));
  A(dalloc(&c->d_gtab, g * g));
  A(dalloc(&c->d_kd, g));
  A(dalloc(&c->d_grids, cells));
  A(dalloc(&c->d_force, cells));
  if (e != hipSuccess) {
    mog2d_release(c);
    delete c;
    return fail(nullptr, MOG_ENOMEM, "device allocation failed for the 2D grids");
  }
  // FP64-derived tables: twiddles, -2 pi G / k^2 / (1000 g^2) with the
  // Nyquist-positive wave numbers (0 at DC), derivative factors with the
  // Nyquist entry zeroed
  std::vector<float2> tw(g);
  std::vector<float> gtab(g * g), kd(g);
  auto sh = [&](int i) { return i <= (int)g / 2 ? i : i - (int)g; };
  for (size_t j = 0; j < g; ++j) {
    const double th = 2.0 * M_PI * (double)j / (double)g;
    tw[j] = make_float2((float)cos(th), (float)sin(th));
    kd[j] = (j == g / 2 && g > 1) ? 0.0f : (float)((double)sh((int)j) * 2.0 * M_PI / world_size);
  }
  for (size_t iy = 0; iy < g; ++iy)
    for (size_t ix = 0; ix < g; ++ix) {
      const double kx = sh((int)ix) * 2.0 * M_PI / world_size, ky = sh((int)iy) * 2.0 * M_PI / world_size;
      const double k2 = kx * kx + ky * ky;
      gtab[iy * g + ix] = (ix | iy) ? (float)(-2.0 * M_PI * G / k2 / (1000.0 * (double)(g * g))) : 0.0f;
    }
  // fast rows and margin of the 3D path: the planar rows have zero z
  // coefficients, and the 3D bound covers the shorter 2D sequence
  const float halff = (float)world_size * 0.5f, hf = (float)(world_size / (double)grid_size);
  std::vector<float> frot((size_t)n_grids * 12);
  double rmax = 0.0, omax = 0.0;
  for (int i = 0; i < n_grids; ++i) {
    make_fast_row(rot_k12 + 12 * i, off_k4 + 4 * i, halff, hf, frot.data() + 12 * i);
    for (int j = 0; j < 12; ++j) rmax = fmax(rmax, fabs((double)rot_k12[12 * i + j]));
    for (int m = 0; m < 2; ++m) omax = fmax(omax, fabs((double)off_k4[4 * i + m]));
  }
  const double margin = fast_margin(world_size, grid_size, rmax, omax, 1.5 * world_size);
  c->lim = margin < 0.05 ? (float)(0.5 - margin) : -1.0f;
  A(hipMemcpy(c->d_rot, rot_k12, (size_t)n_grids * 12 * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(c->d_off, off_k4, (size_t)n_grids * 4 * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(c->d_frot, frot.data(), frot.size() * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(c->d_tw, tw.data(), g * sizeof(float2), hipMemcpyHostToDevice));
  A(hipMemcpy(c->d_gtab, gtab.data(), g * g * sizeof(float), hipMemcpyHostToDevice));
  A(hipMemcpy(c->d_kd, kd.data(), g * sizeof(float), hipMemcpyHostToDevice));
  if (e != hipSuccess) {
    mog2d_release(c);
    delete c;
    return fail(nullptr, MOG_EHIP, "2D tables: %s", hipGetErrorString(e));
  }
  *out = c;
  return MOG_OK;
}

void mog2d_destroy(mog2d_ctx* ctx) {
  if (!ctx) return;
  mog2d_release(ctx);
  delete ctx;
}

int mog2d_forces(mog2d_ctx* ctx, const float* pos_n4, int64_t n, float* avg_n4, void* stream) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (n > 0 && (!pos_n4 || !avg_n4 || !aligned16(pos_n4) || !aligned16(avg_n4)))
    return fail(nullptr, MOG_EINVAL, "pos/avg must be 16-byte aligned device pointers");
  hipError_t e = launch_mog2d_forces((const float4*)pos_n4, n, ctx->k, ctx->lg, (float)ctx->W,
                                     ctx->d_rot, ctx->d_off, ctx->d_frot, ctx->lim, ctx->d_tw,
                                     ctx->d_gtab, ctx->d_kd, ctx->d_grids, ctx->d_force,
                                     (float4*)avg_n4, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog2d_forces: %s", hipGetErrorString(e));
}

int mog2d_deposit(mog2d_ctx* ctx, const float* pos_n4, int64_t n, int64_t* grids_out, void* stream) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (!grids_out) return fail(nullptr, MOG_EINVAL, "grids_out is NULL");
  if (n > 0 && (!pos_n4 || !aligned16(pos_n4)))
    return fail(nullptr, MOG_EINVAL, "pos must be a 16-byte aligned device pointer");
  hipError_t e = launch_mog2d_deposit((const float4*)pos_n4, n, ctx->k, ctx->lg, (float)ctx->W,
                                      ctx->d_rot, ctx->d_off, ctx->d_frot, ctx->lim,
                                      (long long*)grids_out, (hipStream_t)stream);
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog2d_deposit: %s", hipGetErrorString(e));
}

int mog2d_step(mog2d_ctx* ctx, float* pos_n4, float* vel_n4, int64_t n, double dt, double damping,
               void* stream) {
  if (!ctx) return fail(nullptr, MOG_EINVAL, "ctx is NULL");
  if (n < 0 || n > (int64_t)2147483647) return fail(nullptr, MOG_EINVAL, "n must be in [0, 2^31)");
  if (!isfinite(dt) || !isfinite(damping)) return fail(nullptr, MOG_EINVAL, "dt/damping must be finite");
  if (n == 0) return MOG_OK;
  if (!pos_n4 || !vel_n4 || !aligned16(pos_n4) || !aligned16(vel_n4))
    return fail(nullptr, MOG_EINVAL, "pos/vel must be 16-byte aligned device pointers");
  hipStream_t s = (hipStream_t)stream;
  if (ctx->avg_cap < n) {
    (void)hipFree(ctx->d_avg);
    ctx->d_avg = nullptr;
    ctx->avg_cap = 0;
    if (dalloc(&ctx->d_avg, (size_t)n) != hipSuccess)
      return fail(nullptr, MOG_ENOMEM, "2D force buffer for %lld particles", (long long)n);
    ctx->avg_cap = n;
  }
  hipError_t e = launch_mog2d_forces((const float4*)pos_n4, n, ctx->k, ctx->lg, (float)ctx->W,
                                     ctx->d_rot, ctx->d_off, ctx->d_frot, ctx->lim, ctx->d_tw,
                                     ctx->d_gtab, ctx->d_kd, ctx->d_grids, ctx->d_force, ctx->d_avg,
                                     s);
  if (e == hipSuccess) {
    Params P{};
    P.W = (float)ctx->W;
    P.dtp = (float)dt;
    P.damping = (float)damping;
    e = launch_update(P, (float4*)pos_n4, (float4*)vel_n4, ctx->d_avg, n, s);
  }
  return e == hipSuccess ? MOG_OK : fail(nullptr, MOG_EHIP, "mog2d_step: %s", hipGetErrorString(e));
}

const char* mog_last_error(const mog_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

}  // extern "C"
