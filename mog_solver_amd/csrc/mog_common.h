// Shared device helpers and launch parameters of the MOG force step.
//
// The whole library is compiled with -ffp-contract=off: the cell-index
// arithmetic must round every product and sum separately (same order as the
// oracle's float32 emulation) so that density grids are bit-exact.  Code
// that wants fused multiply-adds writes fmaf() explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cellmath.h"
#include "tiling.h"

namespace mog {

// Scalar parameters, float32 exactly as the reference shaders receive them
// (main.py:604-605, 662-667): W = float(world_size), half = W * 0.5f,
// h = float(world_size / grid_size), dtp = float(dt * timescale).
struct Params {
  int k;       // grids
  int g;       // cells per axis (power of two)
  int lg;      // log2(g)
  int cells;   // g^3
  float W, half, h;
  float dtp, damping;
  float fk;    // float(k), the divisor of update.glsl:81
  float CB;    // fast-path admission bound on |c_j| (W / 2: every wrapped position)
  float lim;   // fast-path acceptance: |frac(v) - 0.5| < lim (0.5 - margin); < 0 disables
  int shift_ok;  // kShiftLds | kShiftF4: shifted cell fields valid (cellmath.h field_shift_ok)
};

// Params::shift_ok bits: the shifted fields hold for byte offsets of 4-byte
// (LDS counters, s0 = 2) and 16-byte (float4 forces, s0 = 4) elements of
// rint(w), and of 128-byte elements (s0 = 7) of floor(w) (|floor(w)| <=
// |rint(w)| + 1: the closure's interpolation blocks).
constexpr int kShiftLds = 1;
constexpr int kShiftF4 = 2;
constexpr int kShiftBlk = 4;

// shift_ok for transforms with |R| <= rmax and |off| <= omax.
inline int shift_flags(double W, int g, int lg, double rmax, double omax, double CB) {
  const double wb = fast_w_bound(W, g, rmax, omax, CB);
  return (field_shift_ok(lg, 2, wb) ? kShiftLds : 0) | (field_shift_ok(lg, 4, wb) ? kShiftF4 : 0) |
         (field_shift_ok(lg, 7, wb + 0.5) ? kShiftBlk : 0);
}

// The exact float32 sequence, kept out of line (one copy per translation
// unit): it runs for a few percent of (particle, grid) pairs only.
static __device__ __attribute__((noinline)) int cell_exact_call(float c0, float c1, float c2,
                                                                const float* __restrict__ R,
                                                                const float* __restrict__ O,
                                                                float W, float half, float h,
                                                                int lg) {
  return cell_exact(c0, c1, c2, R, O, W, half, h, lg);
}

// A wave-uniform value copied into a vector register.  A VOP3 instruction
// reads at most one scalar operand on this target, so an FMA whose
// multiplier and addend are both wave-uniform would otherwise re-copy the
// addend with a v_mov on every use.  And on gfx950 a vector instruction
// with a scalar-register operand issues at half the rate of one whose
// operands are vector registers or literals (about 4.3 against 2.7 cycles
// per wave64 add, FMA or bit operation, measured), so the hot loops keep
// their wave-uniform coefficients in vector registers.
__device__ __forceinline__ float to_vgpr(float x) {
  float y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
  return y;
}

__device__ __forceinline__ unsigned to_vgpr(unsigned x) {
  unsigned y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
  return y;
}

// Cell of the centred position (c0,c1,c2) in one grid: the fast path when
// the particle is admitted (pfast) and v is clear of every cell face,
// otherwise the exact float32 sequence of the reference (cellmath.h).
// F points at the grid's fast row (registers or scalar memory), R and O at
// its packed rotation row and offset in global memory.
__device__ __forceinline__ int cell_index(float c0, float c1, float c2, int pfast, const float* F,
                                          const float* __restrict__ R,
                                          const float* __restrict__ O, const Params& P) {
  int cell;
  const int ok = cell_fast(c0, c1, c2, F, P.lim, P.lg, &cell) & pfast;
  if (!ok) cell = cell_exact_call(c0, c1, c2, R, O, P.W, P.half, P.h, P.lg);
  return cell;
}

// Cell `cell` (0 <= cell < 2^28) of a grid of float4: a 32-bit byte offset
// on a wave-uniform base, so the load takes the scalar-base addressing form.
__device__ __forceinline__ float4 cell_load(const float4* __restrict__ grid, int cell) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(grid) +
                                          ((unsigned)cell << 4));
}

// The float4 at byte offset `off` (< 2^32) of a grid: wave-uniform base,
// 32-bit vector offset.
__device__ __forceinline__ float4 cell_load_b(const float4* __restrict__ grid, unsigned off) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(grid) + off);
}

// Position of the cached cell of (slot t, grid) in the uint16 cell cache:
// tiles of 64 slots (one wave) x k grids, so a wave's deposit writes and its
// gather reads whole contiguous rows.
__device__ __forceinline__ long long cache_index(long long t, int grid, int k) {
  return (((t >> 6) * k + grid) << 6) | (t & 63);
}

// uint(mass * 1000.0) of mass.glsl:54 as a 64-bit integer; negative and NaN
// masses deposit 0.
__device__ __forceinline__ unsigned long long milli_of(float m) {
  float mm = m * 1000.0f;
  return mm > 0.0f ? (unsigned long long)mm : 0ull;
}

// The integrator of update.glsl:83-86, each operation rounded in order:
// vel -= avgForce * deltaTime; vel *= damping; pos += vel * deltaTime;
// pos = mod(pos, worldSize); stores (pos, m) and (vel, 0).
__device__ __forceinline__ void integrate(float4& p, float4& v, float ax, float ay, float az,
                                          const Params& P) {
  float vx = v.x - ax * P.dtp;
  float vy = v.y - ay * P.dtp;
  float vz = v.z - az * P.dtp;
  vx = vx * P.damping;
  vy = vy * P.damping;
  vz = vz * P.damping;
  const float x = p.x + vx * P.dtp;
  const float y = p.y + vy * P.dtp;
  const float z = p.z + vz * P.dtp;
  p = make_float4(glsl_mod_f(x, P.W), glsl_mod_f(y, P.W), glsl_mod_f(z, P.W), p.w);
  v = make_float4(vx, vy, vz, 0.0f);
}

}  // namespace mog

// Host-side launchers (defined in the kernel translation units).
namespace mog {

// Rows of the prepared deposit: (c0, c1, c2, milli) with c = p - W/2 and the
// 32-bit milli-mass, or kPrepMarker where the fast path cannot take the
// particle (outside |c_j| <= CB, non-finite, or a milli-mass above the
// partials' limit).  The milli-mass and the admission test are those of the
// unprepared loop (same float32 operations, saturating conversion).  Written
// by the deposit's prepass, or by a run's gather for the next step.
constexpr unsigned int kPrepMarker = 0xFFFFFFFFu;

__device__ __forceinline__ float4 prepared_row(const float4 q, const Params& P,
                                               unsigned int lds_limit32) {
  const float c0 = q.x - P.half, c1 = q.y - P.half, c2 = q.z - P.half;
  const float mm = q.w * 1000.0f;
  unsigned int m;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(m) : "v"(mm));
  const bool take = (m <= lds_limit32) & (bool)particle_fast(c0, c1, c2, P.CB);
  return make_float4(c0, c1, c2, __uint_as_float(take ? m : kPrepMarker));
}

struct DepositTune {
  int threads;        // threads per workgroup
  int grids_per_wg;   // grids privatised in LDS per workgroup
  int wg_target;      // aim for about this many workgroups
  int prepass = 0;    // prepared rows: 0 from 2^26 pairs, 1 always, -1 never
  long long min_chunk = 0;  // particles per workgroup at least this many
  bool cpos_ready = false;  // the prepared rows are already in cpos: no prepass
};

hipError_t launch_fill_zero(long long* p, long long count, hipStream_t s);
// cpos (optional, n rows): scratch for the prepared rows of the LDS
// deposit (centred coordinates and milli-mass, written by a prepass).
hipError_t launch_deposit(const Params& P, const float4* pos, long long n, const float* rot,
                          const float* off, const float* frot, unsigned long long* grids,
                          const DepositTune& t, hipStream_t s, float4* cpos = nullptr,
                          unsigned int* prep_limit = nullptr);
// prep_limit (optional): set to the prepared rows' milli-mass limit when the
// deposit reads prepared rows (with t.cpos_ready it then runs no prepass),
// else to 0.  launch_deposit_prep writes the rows of n particles.
hipError_t launch_deposit_prep(const Params& P, const float4* pos, long long n,
                               unsigned int lds_limit32, float4* cpos, hipStream_t s);
// world != 0: force grids hold M F (world frame) for the step's gather;
// world == 0: grid-frame forces as pack.glsl writes them.
// Deposit of positions already in Morton visiting order (sorted != 0; lanes
// aggregate shared cells) with an optional uint16 cell cache
// cells_out[cache_index(slot, grid, k)] (g <= 32 only, sorted only).
hipError_t launch_deposit_ex(const Params& P, const float4* pos, long long n, const float* rot,
                             const float* off, const float* frot, unsigned long long* grids,
                             unsigned short* cells_out, bool sorted, const DepositTune& t,
                             hipStream_t s, float4* cpos = nullptr,
                             unsigned int* prep_limit = nullptr);
// spos[t] = pos[order[t]]
hipError_t launch_permute(const float4* pos, const int* order, long long n, float4* spos,
                          hipStream_t s);
// Gather + update of Morton-ordered slots from the cell cache: slot t is
// particle order[t] and its cell in grid g is cells[cache_index(t, g, k)].
hipError_t launch_gather_cached(const Params& P, float4* pos, float4* vel, long long n,
                                const int* order, const unsigned short* cells,
                                const float4* force, hipStream_t s);
// conv != NULL: multiply by the spectrum of a real-space kernel instead of
// the Green's function x i*k (conv[0..g^3) = T_y + i T_z, conv[g^3..) = T_x).
hipError_t launch_solve(const Params& P, const long long* grids, float2* S1, long long s1_elems,
                        float2* S2, float4* force, const float2* tw, const float* gtab,
                        const float* kd, const float2* conv, int need_x, const float* rot,
                        int world, hipStream_t s);
// float2 elements of S1 the g = 32 solve needs for `ngrids` grids (one
// scratch slot per workgroup, at most one workgroup per CU)
long long rh_scratch_elems(long long ngrids);
// Entries of the convolution solver's multiplier table of one simulation:
// T_y + i T_z and T_x over the full spectrum (2 g^3), then for g = 32 the
// three half-spectrum multipliers T_x, T_y, T_z (kx <= g/2) of the g = 32
// kernel.
inline long long conv_table_elems(int g) {
  const long long g3 = (long long)g * g * g;
  return 2 * g3 + (g == 32 ? 3LL * g * g * (g / 2 + 1) : 0);
}
// MFMA spectral solve (g >= 16): precision 0 FP32, 1 BF16X3, 2 FP16.
// Wf / Wi: forward / inverse DFT tables, each [Re W | -Im W] then
// [Im W | Re W] as g rows of 2g floats.
hipError_t launch_solve_mfma(const Params& P, int precision, const long long* grids, float2* S1,
                             float2* S2, float4* force, const float* Wf, const float* Wi,
                             const float* gtab, const float* kd, const float2* conv, int need_x,
                             const float* rot, int world, hipStream_t s);
// zero != NULL: the launch also clears zero[0, zero_count) (the step's
// spare density grids), spread over all of its lanes.
// ps, vs (optional, with order): the particle state in visiting order,
// ps[t] = pos[order[t]], vs[t] = vel[order[t]]; the gather then reads and
// writes them at slot t (coalesced) and writes only the positions back to
// pos[order[t]] (the deposit's input), leaving vel stale until
// launch_state_export.  cpos_out (optional, with ps): the next deposit's
// prepared row goes to cpos_out[order[t]] instead (milli-mass limit
// lds_limit32), and pos stays stale too.
hipError_t launch_gather_update(const Params& P, float4* pos, float4* vel, long long n,
                                const float* rot, const float* off, const float* frot,
                                const float4* force, int world, float4* avg_out, int mode,
                                const int* order, hipStream_t s, long long* zero = nullptr,
                                long long zero_count = 0, float4* ps = nullptr,
                                float4* vs = nullptr, float4* cpos_out = nullptr,
                                unsigned int lds_limit32 = 0);
// ps[t] = pos[order[t]], vs[t] = vel[order[t]] for t < n.
hipError_t launch_state_import(const float4* pos, const float4* vel, const int* order, long long n,
                               float4* ps, float4* vs, hipStream_t s);
// vel[order[t]] = vs[t] for t < n (and pos[order[t]] = ps[t] when ps is
// given: a run whose gather wrote prepared rows instead of positions).
hipError_t launch_state_export(const float4* vs, const int* order, long long n, float4* vel,
                               hipStream_t s, const float4* ps = nullptr, float4* pos = nullptr);
// Morton visiting order of the particles (order.hip): temp storage size of
// the radix sort, and the order itself (keys_a/keys_b/idx_a: n-element
// scratch; order: n-element output).
hipError_t order_temp_bytes(long long n, size_t* bytes);
hipError_t build_order(const float4* pos, long long n, float W, unsigned* keys_a,
                       unsigned* keys_b, int* idx_a, int* order, void* tmp, size_t tmp_bytes,
                       hipStream_t s);
hipError_t order_temp_bytes_db(long long n, size_t* bytes);
hipError_t build_order_db(const float4* pos, long long n, float W, unsigned* keys_a,
                          unsigned* keys_b, int* idx_a, int* idx_b, void* tmp, size_t tmp_bytes,
                          const int** order, hipStream_t s);
// Device-resident scratch of the no-allocation step (noalloc.hip): a
// file-scope __device__ array of kNoAllocBytes per GPU, holding the visiting
// order of up to kNoAllocMaxParticles particles (four n-element 32-bit
// arrays) followed by at most kNoAllocSortBytes of sort temporaries.
// noalloc_scratch returns its device address on the current device (looked
// up once per device with hipGetSymbolAddress).
constexpr long long kNoAllocMaxParticles = 1ll << 20;
constexpr size_t kNoAllocSortBytes = (size_t)4 << 20;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocMaxParticles * 16 + kNoAllocSortBytes;
hipError_t noalloc_scratch(void** ptr);
hipError_t launch_init_perlin(float4* pos, float4* vel, long long n, float W, float noise_scale,
                              float contrast, unsigned seed, float base_mass, float spawn_buffer,
                              int sphere, hipStream_t s);
hipError_t launch_solve_pair(const Params& P, const long long* grids, float4* force_a,
                             const float* gtab_a, const float2* conv_a, float4* force_b,
                             const float* gtab_b, const float2* conv_b, const float* kd,
                             int need_x, const float* rot, int world, hipStream_t s);
// Closure features: a per-cell pass over the k context grids writes fm
// (k * cells float4) and blk (k * cells * 8 float4), then the per-particle
// pass reads them.
hipError_t launch_closure_features(const Params& P, const float4* pos, long long n,
                                   const float* rot, const float* off, const float* frot,
                                   const float4* force, const float4* force_dc,
                                   const long long* grids, float rho0, const int* order,
                                   float4* fm, float4* blk, float4* feat, float4* mean,
                                   hipStream_t s);
size_t mlp_param_count();
hipError_t launch_direct_sum(const float4* pos, long long n, const int* targets, int nt, double W,
                             double G_eff, double eps2, double* out, hipStream_t s);
hipError_t launch_mlp_forward(const float* prm, const float4* feat, const float4* mean,
                              long long n, float4* out, bool bf16, hipStream_t s);
hipError_t launch_mlp_train(const float* prm, const float4* feat, const float4* mean,
                            const float4* target, long long n, float wnorm, float* grad,
                            float* loss, bool bf16, hipStream_t s);
hipError_t launch_adam(float* prm, const float* grad, float* m1, float* m2, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s);
hipError_t launch_update(const Params& P, float4* pos, float4* vel, const float4* avg,
                         long long n, hipStream_t s);
// FP64 diagnostics (energy.hip): out[0..7] = (sum m v (3), 1/2 sum m |v|^2,
// mean grid potential energy, sum m, 0, 0) of the particles and the k int64
// density grids; tw[j] = exp(-2 pi i j / g), tab = T(k) = G_hat(k) /
// (2 g^3 10^6) as (g,g,g) [kz][ky][kx]; scratch: diag_scratch_doubles(k, g).
size_t diag_scratch_doubles(int k, int g);
hipError_t launch_diagnostics(const float4* pos, const float4* vel, long long n,
                              const long long* grids, int k, int g, const double2* tw,
                              const double* tab, double* scratch, double* out, hipStream_t s);


// ---------------------------------------------------------------------------
// Batched ensemble: many independent simulations in packed arrays
// ---------------------------------------------------------------------------
// One simulation of the ensemble and where its data live.  Particles of sim
// s are rows [p_off, p_off + n) of the packed pos/vel; its k grids are cells
// [cell_off, cell_off + k g^3) of the packed grids / S1 / S2 / force; its
// transforms are rows [grid_off, grid_off + k) of rot / off / frot; its
// solve slabs (k g per pass) are [slab_off, slab_off + k g) of its
// grid-size class's flat slab range.
struct BatchSim {
  Params P;
  long long p_off, n;
  long long cell_off, slab_off;
  long long conv_off;  // float2 offset of its convolution spectrum; < 0: Green's function
  long long blk_off;   // first gather block (ceil(n / kBatchGatherBlock) blocks per sim)
  int grid_off;
  int tab_off;         // gtab offset (g^3 floats)
  int line_off;        // tw / kd offset (g entries)
  int need_x;
};

// A deposit work item: particles [p0, p1) of `sim` into its grids
// [grid0, grid0 + ng), privatised in LDS.
struct DepositTask {
  int sim, grid0, ng, G;  // G: the deposit tile's histogram count (1, 2 or 4), ng <= G
  long long p0, p1;
};

// Largest LDS histogram of a batched deposit task (uint32 counters).
constexpr int kBatchLdsCells = 32768;
// Most grids per task of the batched deposit (fast rows in vector registers).
constexpr int kBatchMaxG = 4;
int batch_task_grids(int cells);
// Particles per gather block; a block never straddles two simulations.
constexpr int kBatchGatherBlock = 256;
// Deferred undecided pairs per lane of a gather (grid indices in LDS): the
// exact cells of a lane's first kGatherSlots undecided pairs are evaluated
// after its grid loop, the rest on the spot.
constexpr int kGatherSlots = 4;

// The simulation owning gather block `blk` (sims ordered by blk_off); the
// result is uniform across a workgroup that passes a uniform blk.
__device__ __forceinline__ int sim_of_block(const BatchSim* __restrict__ sims, int nsims,
                                            long long blk) {
  int lo = 0, hi = nsims - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sims[mid].blk_off <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Visiting order of the batch gather: particles of each simulation sorted
// by Morton key inside the simulation's own rows (one radix sort of
// (sim << 30 | key) over all particles).
hipError_t batch_order_temp_bytes(long long n, int nsims, size_t* bytes);
hipError_t build_batch_order(const float4* pos, long long n, long long blocks,
                             const BatchSim* sims, int nsims, unsigned long long* keys_a,
                             unsigned long long* keys_b, int* idx_a, int* order, void* tmp,
                             size_t tmp_bytes, hipStream_t s);

// 2D MOG: clear, deposit, per-grid 2D spectral solve and gather of the
// average world force (avg = (Fx, Fy, 0, 0)).  grids: k g^2 int64 scratch,
// force: k g^2 float2 scratch; tw / gtab / kd: the 2D tables.
// frot: the grids' fast rows (make_fast_row), lim: 0.5 - margin (< 0
// disables the fast path).
hipError_t launch_mog2d_forces(const float4* pos, long long n, int k, int lg, float W,
                               const float* rot, const float* off, const float* frot, float lim,
                               const float2* tw, const float* gtab, const float* kd,
                               long long* grids, float2* force, float4* avg, hipStream_t s);
// clear + deposit only: the k g^2 int64 density grids of the 2D path.
hipError_t launch_mog2d_deposit(const float4* pos, long long n, int k, int lg, float W,
                                const float* rot, const float* off, const float* frot, float lim,
                                long long* grids, hipStream_t s);

// The batch's prepass (prepared rows into cpos, one row per particle; the
// partials' limit of a task of `chunk` particles) and its persistent deposit
// over the task list; blocks: the batch's gather blocks (kBatchGatherBlock
// particles of one simulation each).
hipError_t launch_batch_deposit(const float4* pos, float4* cpos, const BatchSim* sims, int nsims,
                                long long blocks, long long chunk, const DepositTask* tasks,
                                int ntasks, const float* rot, const float* off, const float* frot,
                                unsigned long long* grids, unsigned long long* counter,
                                int threads, hipStream_t s);
// Solve passes over one window of a grid-size class: simulations
// cls_sim[0..ncls) whose slabs start at cls_off[j] in the class's slab
// range; the window is slabs [slab0, slab0 + slabs) (whole grids) and the
// spectral scratch S1/S2 of class slot j starts at offset soff[j] (complex
// elements; see kBatchWindowCells).  LDS is sized for the class's largest
// grid `gmax`.
hipError_t launch_batch_solve(const BatchSim* sims, const int* cls_sim, const long long* cls_off,
                              int ncls, long long slab0, long long slabs, const long long* soff,
                              int gmax, const long long* grids, float2* S1, long long s1_elems,
                              float2* S2, float4* force, const float2* tw, const float* gtab,
                              const float* kd, const float2* conv, const float* rot,
                              hipStream_t s);
// Cells of the batch's spectral scratch window for the grid sizes that run
// the slab passes (g <= 8): their solve runs window by window through one
// reused pair of complex buffers, which so stay in the Infinity Cache
// instead of streaming every simulation's scratch through HBM.
constexpr long long kBatchWindowCells = 1ll << 22;
// Grid-size classes whose solve is one launch over all of the class's
// grids (one workgroup per grid, no windows).
inline bool batch_class_fused(int gmax) { return gmax == 16 || gmax == 32; }
// ps, vs (optional, with order): the ensemble's state in visiting order,
// as for launch_gather_update.
hipError_t launch_batch_gather(float4* pos, float4* vel, long long blocks, const BatchSim* sims,
                               int nsims, int kmax, const float* rot, const float* off,
                               const float* frot, const float4* force, const int* order,
                               hipStream_t s, float4* ps = nullptr, float4* vs = nullptr);
}  // namespace mog
