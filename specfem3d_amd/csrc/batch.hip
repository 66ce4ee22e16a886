// batch.hip -- batched many-shot forward engine: many independent shots (own sources, receivers
// and record lengths) on one mesh prepared through the reference C-ABI.
//
// The reference runs one shot per process (one Mesh_pointer, one source set, iterate_time.F90).
// Here a Batch borrows a prepared elastic Mesh (ibool, moduli, Jacobians, inverse mass, Stacey
// face points, element colours) and steps S interleaved wavefields together:
//   * layout: value (g, c, s) at ((int64_t)g * 3 + c) * S + s (shot-fastest rows, see k_batch.hip);
//   * shots are placed in decreasing order of record length, so the live shots at any step are
//     a prefix of the row and finished shots cost no memory traffic;
//   * a mesh with fluid elements also steps S interleaved potentials over its acoustic range of
//     global points (value (g, s) at (g - g0) * S + s): per step the potential predictor, free
//     surface, acoustic stiffness, fluid Stacey and fluid side of the coupling, the corrector and
//     the free surface again come first, then the solid side of the coupling with the elastic
//     boundary terms (the single-shot call order of iterate_time.F90);
//   * one step = fused corrector/predictor, Stacey rounds, source rounds, stiffness (one launch per
//     element colour with plain read-modify-writes, or one atomic launch), receivers, and a
//     one-lane kernel that advances the device step counter. Assembly mode 3 instead records the
//     receivers first and folds the corrector/predictor into the coloured stiffness launches
//     (k_batch.hip), which removes the separate pass over u, v, a. Every kernel reads the counter,
//     so a group of steps is captured once into a hipGraph and replayed;
//   * the per-point arithmetic and the order of every sum equal the single-shot path (same
//     kernels' expressions, same colour order, same face rounds), so each shot of a batch is
//     bit-identical to its single-shot run with the same options.
// Shots-per-workgroup and the assembly mode are chosen by an autotuner that times each candidate
// for a few steps from rest and then resets the batch.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "kernels.h"

namespace sf {
void ensure_element_colors(Mesh* m);
void ensure_acoustic_colors(Mesh* m);
bool element_is_elastic(Mesh* m, int e);

struct Batch {
  Mesh* m = nullptr;
  int nshots = 0, S = 0;
  int S0 = 0;                    // padded shot count at creation; S shrinks as shots finish
  int sc = 4, mode = 2;          // shots per workgroup, assembly (0 atomic, 2 coloured, 3 fused)
  int overlay = 0;               // 1: displacement tile and stress products share LDS
  int max_nstep = 0;
  realw dt = 0.f, dt2 = 0.f, dtsq2 = 0.f;  // deltat, deltatover2, deltatsqover2
  std::vector<int> nstep_user;   // per user shot
  std::vector<int> slot_of;      // user shot -> slot
  int* d_nstep = nullptr;        // per slot (S entries, decreasing, padding 0)
  int* d_it = nullptr;
  int h_it = 1;                  // host mirror of the step counter
  realw *u = nullptr, *v = nullptr, *a = nullptr;
  realw* u_last = nullptr;       // final displacement of slots dropped by compaction (stride S0)
  std::vector<int> h_nstep;      // per slot (S0 entries, decreasing, padding 0)
  bool att = false;              // 3-SLS attenuation: per-shot memory variables and strain
  bool ac = false;               // fluid elements: per-shot potentials over the acoustic range
  int64_t ac_g0 = 0, ac_n = 0;   // acoustic range [ac_g0, ac_g0 + ac_n) of global points
  realw *pot = nullptr, *potd = nullptr, *potdd = nullptr;
  realw* p_last = nullptr;       // final potential of slots dropped by compaction (stride S0)
  realw *R = nullptr, *eps = nullptr;
  unsigned char* flags = nullptr;  // fused mode: BF_* per local point (first/last colour, cleared)
  int flags_version = -1;        // Mesh::color_version the flags were built for (-1: stale)
  // sources: (source, point) entries sorted into rounds (the k-th source of each shot)
  struct Src { int shot, ispec; std::vector<realw> arr, stf; };
  std::vector<Src> srcs;
  int src_n = 0, src_nrounds = 0;
  std::vector<int> src_round_off;
  int *src_ig = nullptr, *src_slot = nullptr, *src_id = nullptr, *src_ploc = nullptr;
  realw *src_arr = nullptr, *src_stf = nullptr;
  // receivers
  struct Rec { int shot, ispec; double h[125]; double nu[9]; };
  std::vector<Rec> recs;
  int nrec = 0;
  std::vector<int> rec_order;    // device receiver index -> position in recs
  int *rec_slot = nullptr, *rec_ig = nullptr;
  double *rec_h = nullptr, *rec_nu = nullptr;
  realw* seis = nullptr;         // [max_nstep][nrec][3]
  bool ready = false;
  // captured step groups
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  // timing of the stiffness launches (sf_batch_time)
  bool timing = false;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
};

static Batch* batch_of(long h) {
  if (!h) fatal("sf_batch", "null batch handle");
  return reinterpret_cast<Batch*>(h);
}

// ------------------------------------------------------------------------------ kernels
static __device__ __forceinline__ int load_it(const int* it) {
  return __hip_atomic_load(it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Newmark update of every live (row, shot): predictor only on the first step, otherwise the
// deferred corrector of the previous step fused with the predictor (k_update.hip arithmetic).
// A lane takes four consecutive shots of one row (a float4); a workgroup takes 256 / (S/4) rows,
// so the row and quad of a lane come from 32-bit arithmetic on threadIdx. Finished shots keep
// their values (their lanes store what they loaded).
__global__ void __launch_bounds__(256) k_batch_update(int64_t nrow, int S, int rows_per_block,
                                                      const int* __restrict__ nstep, const int* __restrict__ itp,
                                                      realw* __restrict__ u, realw* __restrict__ v,
                                                      realw* __restrict__ a, const realw* __restrict__ rm, realw dt2p,
                                                      realw dt, realw dt2, realw dtsq2, int first) {
  const int nq = S >> 2;
  const int rl = (int)threadIdx.x / nq;
  const int qd = (int)threadIdx.x - rl * nq;
  if (rl >= rows_per_block) return;
  const int64_t row = (int64_t)blockIdx.x * rows_per_block + rl;
  if (row >= nrow) return;
  const int it = load_it(itp);
  const int s = 4 * qd;
  if (nstep[s] < it) return;                       // shots are sorted: the whole quad has finished
  const bool l1 = nstep[s + 1] >= it, l2 = nstep[s + 2] >= it, l3 = nstep[s + 3] >= it;
  const int64_t o = row * S + s;
  const float4 U0 = *reinterpret_cast<const float4*>(u + o);
  const float4 V0 = *reinterpret_cast<const float4*>(v + o);
  const float4 A0 = *reinterpret_cast<const float4*>(a + o);
  const realw M = first ? (realw)0 : rm[row];
  float4 U = U0, V = V0, A = A0;
#define SF_BU(c)                           \
  if (!first) {                            \
    A.c = A.c * M;                         \
    V.c = V.c + dt2p * A.c;                \
  }                                        \
  U.c = U.c + dt * V.c + dtsq2 * A.c;      \
  V.c = V.c + dt2 * A.c;                   \
  A.c = 0.f;
  SF_BU(x) SF_BU(y) SF_BU(z) SF_BU(w)
#undef SF_BU
  if (!l1) { U.y = U0.y; V.y = V0.y; A.y = A0.y; }
  if (!l2) { U.z = U0.z; V.z = V0.z; A.z = A0.z; }
  if (!l3) { U.w = U0.w; V.w = V0.w; A.w = A0.w; }
  *reinterpret_cast<float4*>(u + o) = U;
  *reinterpret_cast<float4*>(v + o) = V;
  *reinterpret_cast<float4*>(a + o) = A;
}

// Stacey (k_stacey_el arithmetic) for every (face point of one round, live shot)
__global__ void k_batch_stacey(int n, int S, const int* __restrict__ ig, const realw* __restrict__ nrm,
                               const realw* __restrict__ jw, const realw* __restrict__ rvp,
                               const realw* __restrict__ rvs, const int* __restrict__ nstep,
                               const int* __restrict__ itp, const realw* __restrict__ veloc, realw* accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // n * S < 2^31 (checked by the caller)
  const int q = t / S;
  const int s = t - q * S;
  if (q >= n) return;
  if (nstep[s] < load_it(itp)) return;
  const int64_t b = (int64_t)ig[q] * 3 * S + s;
  const realw vx = veloc[b], vy = veloc[b + S], vz = veloc[b + 2 * S];
  const realw nx = nrm[3 * q], ny = nrm[3 * q + 1], nz = nrm[3 * q + 2];
  const realw vn = vx * nx + vy * ny + vz * nz;
  const realw p = rvp[q], sv = rvs[q];
  const realw tx = p * vn * nx + sv * (vx - vn * nx);
  const realw ty = p * vn * ny + sv * (vy - vn * ny);
  const realw tz = p * vn * nz + sv * (vz - vn * nz);
  const realw w = jw[q];
  const realw fx = tx * w, fy = ty * w, fz = tz * w;
  accel[b] = accel[b] - fx;
  accel[b + S] = accel[b + S] - fy;
  accel[b + 2 * S] = accel[b + 2 * S] - fz;
}

// source entries of one round: accel += sourcearray * stf(it) (k_add_sources_el arithmetic)
__global__ void k_batch_sources(int n, int S, int max_nstep, const int* __restrict__ ig, const int* __restrict__ slot,
                                const int* __restrict__ sid, const int* __restrict__ ploc,
                                const realw* __restrict__ arr, const realw* __restrict__ stf,
                                const int* __restrict__ nstep, const int* __restrict__ itp, realw* accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int s = slot[t];
  const int it = load_it(itp);
  if (nstep[s] < it) return;
  const realw* sa = arr + ((int64_t)sid[t] * 125 + ploc[t]) * 3;
  const realw st = stf[(int64_t)sid[t] * max_nstep + it - 1];
  const int64_t b = (int64_t)ig[t] * 3 * S + s;
  accel[b] = accel[b] + sa[0] * st;
  accel[b + S] = accel[b + S] + sa[1] * st;
  accel[b + 2 * S] = accel[b + 2 * S] + sa[2] * st;
}

// receivers (k_seismo_el arithmetic): one wave64 per receiver, FP64 sums, rotation by nu
__global__ void __launch_bounds__(64) k_batch_seismo(int nrec, int S, const int* __restrict__ slot,
                                                     const int* __restrict__ ig, const double* __restrict__ h,
                                                     const double* __restrict__ nu, const int* __restrict__ nstep,
                                                     const int* __restrict__ itp, const realw* __restrict__ u,
                                                     realw* __restrict__ seis) {
  const int r = blockIdx.x;
  if (r >= nrec) return;
  const int s = slot[r];
  const int it = load_it(itp);
  if (nstep[s] < it) return;
  const int lane = threadIdx.x;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int p = lane; p < 125; p += 64) {
    const int64_t b = (int64_t)ig[(int64_t)r * 125 + p] * 3 * S + s;
    const double hl = h[(int64_t)r * 125 + p];
    sx += (double)u[b] * hl;
    sy += (double)u[b + S] * hl;
    sz += (double)u[b + 2 * S] * hl;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, 64);
    sy += __shfl_xor(sy, off, 64);
    sz += __shfl_xor(sz, off, 64);
  }
  if (lane == 0) {
    const double* n = nu + (int64_t)r * 9;
    realw* o = seis + ((int64_t)(it - 1) * nrec + r) * 3;
    o[0] = (realw)(n[0] * sx + n[1] * sy + n[2] * sz);
    o[1] = (realw)(n[3] * sx + n[4] * sy + n[5] * sz);
    o[2] = (realw)(n[6] * sx + n[7] * sy + n[8] * sz);
  }
}

// ---- fluid domain: potentials over the acoustic range, value (g, s) at (g - g0) * S + s
// Newmark on the potential rows of live shots (k_update.hip arithmetic): correct = 0 the
// predictor chi += dt chi' + dt^2/2 chi'', chi' += dt/2 chi'', chi'' = 0; correct = 1 the corrector
// chi'' *= rmass, chi' += dt/2 chi''. A lane takes four consecutive shots of one row.
__global__ void __launch_bounds__(256) k_batch_ac_newmark(int64_t nrow, int S, int rows_per_block,
                                                          const int* __restrict__ nstep, const int* __restrict__ itp,
                                                          realw* __restrict__ pot, realw* __restrict__ potd,
                                                          realw* __restrict__ potdd, const realw* __restrict__ rm,
                                                          realw dt, realw dt2, realw dtsq2, int correct) {
  const int nq = S >> 2;
  const int rl = (int)threadIdx.x / nq;
  const int qd = (int)threadIdx.x - rl * nq;
  if (rl >= rows_per_block) return;
  const int64_t row = (int64_t)blockIdx.x * rows_per_block + rl;
  if (row >= nrow) return;
  const int it = load_it(itp);
  const int s = 4 * qd;
  if (nstep[s] < it) return;
  const bool l1 = nstep[s + 1] >= it, l2 = nstep[s + 2] >= it, l3 = nstep[s + 3] >= it;
  const int64_t o = row * S + s;
  const float4 V0 = *reinterpret_cast<const float4*>(potd + o);
  const float4 A0 = *reinterpret_cast<const float4*>(potdd + o);
  float4 V = V0, A = A0;
  if (correct) {
    const realw M = rm[row];
#define SF_AC(c) A.c = A.c * M; V.c = V.c + dt2 * A.c;
    SF_AC(x) SF_AC(y) SF_AC(z) SF_AC(w)
#undef SF_AC
    if (!l1) { V.y = V0.y; A.y = A0.y; }
    if (!l2) { V.z = V0.z; A.z = A0.z; }
    if (!l3) { V.w = V0.w; A.w = A0.w; }
    *reinterpret_cast<float4*>(potd + o) = V;
    *reinterpret_cast<float4*>(potdd + o) = A;
    return;
  }
  const float4 U0 = *reinterpret_cast<const float4*>(pot + o);
  float4 U = U0;
#define SF_AP(c) U.c = U.c + dt * V.c + dtsq2 * A.c; V.c = V.c + dt2 * A.c; A.c = 0.f;
  SF_AP(x) SF_AP(y) SF_AP(z) SF_AP(w)
#undef SF_AP
  if (!l1) { U.y = U0.y; V.y = V0.y; A.y = A0.y; }
  if (!l2) { U.z = U0.z; V.z = V0.z; A.z = A0.z; }
  if (!l3) { U.w = U0.w; V.w = V0.w; A.w = A0.w; }
  *reinterpret_cast<float4*>(pot + o) = U;
  *reinterpret_cast<float4*>(potd + o) = V;
  *reinterpret_cast<float4*>(potdd + o) = A;
}

// free surface (k_free_surface): zero chi, chi', chi'' of every (free-surface point, live shot)
__global__ void k_batch_free_surface(int n, int S, int64_t g0, const int* __restrict__ ig,
                                     const int* __restrict__ nstep, const int* __restrict__ itp, realw* pot,
                                     realw* potd, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // n * S < 2^31 (checked by the caller)
  const int q = t / S;
  const int s = t - q * S;
  if (q >= n) return;
  if (nstep[s] < load_it(itp)) return;
  const int64_t b = ((int64_t)ig[q] - g0) * S + s;
  pot[b] = 0.f;
  potd[b] = 0.f;
  potdd[b] = 0.f;
}

// fluid Stacey (k_stacey_ac arithmetic) for every (face point of one round, live shot)
__global__ void k_batch_stacey_ac(int n, int S, int64_t g0, const int* __restrict__ ig, const realw* __restrict__ jw,
                                  const realw* __restrict__ rho, const realw* __restrict__ cp,
                                  const int* __restrict__ nstep, const int* __restrict__ itp,
                                  const realw* __restrict__ potd, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = t / S;
  const int s = t - q * S;
  if (q >= n) return;
  if (nstep[s] < load_it(itp)) return;
  const int64_t b = ((int64_t)ig[q] - g0) * S + s;
  const realw absorbl = potd[b] * jw[q] / cp[q] / rho[q];
  potdd[b] = potdd[b] - absorbl;
}

// fluid side of the coupling (k_coupling_ac_el arithmetic, forward field): chi'' += w (u . n)
__global__ void k_batch_coupling_ac_el(int n, int S, int64_t g0, const int* __restrict__ ig,
                                       const realw* __restrict__ nrm, const realw* __restrict__ jw,
                                       const int* __restrict__ nstep, const int* __restrict__ itp,
                                       const realw* __restrict__ u, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = t / S;
  const int s = t - q * S;
  if (q >= n) return;
  if (nstep[s] < load_it(itp)) return;
  const int64_t g = ig[q];
  const int64_t e = g * 3 * S + s;
  const realw ux = u[e], uy = u[e + S], uz = u[e + 2 * S];
  const realw un = ux * nrm[3 * q] + uy * nrm[3 * q + 1] + uz * nrm[3 * q + 2];
  const int64_t b = (g - g0) * S + s;
  potdd[b] = potdd[b] + jw[q] * un;
}

// solid side of the coupling (k_coupling_el_ac arithmetic, forward field): a += w n (-chi'')
__global__ void k_batch_coupling_el_ac(int n, int S, int64_t g0, const int* __restrict__ ig,
                                       const realw* __restrict__ nrm, const realw* __restrict__ jw,
                                       const int* __restrict__ nstep, const int* __restrict__ itp,
                                       const realw* __restrict__ potdd, realw* accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = t / S;
  const int s = t - q * S;
  if (q >= n) return;
  if (nstep[s] < load_it(itp)) return;
  const int64_t g = ig[q];
  const realw pr = -potdd[(g - g0) * S + s];
  const realw w = jw[q];
  const int64_t e = g * 3 * S + s;
  accel[e] = accel[e] + w * nrm[3 * q] * pr;
  accel[e + S] = accel[e + S] + w * nrm[3 * q + 1] * pr;
  accel[e + 2 * S] = accel[e + 2 * S] + w * nrm[3 * q + 2] * pr;
}

__global__ void k_batch_tick(int* itp) {
  if (threadIdx.x == 0) itp[0] = itp[0] + 1;
}

__global__ void k_batch_fill(int64_t n, realw* p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.f;
}

// Compaction of shot-interleaved rows (grid-stride loops over int64 counts; launched with at most
// 2^20 workgroups): dst[r * S2 + s] = src[r * S + s] for s < S2 (S2 <= S, multiples of 8; float4)
__global__ void k_batch_restride(int64_t nrow, int S, int S2, const realw* __restrict__ src,
                                 realw* __restrict__ dst) {
  const int nq = S2 >> 2;
  const int64_t total = nrow * nq;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / nq;
    const int k = (int)(t - r * nq);
    *reinterpret_cast<float4*>(dst + r * S2 + 4 * k) = *reinterpret_cast<const float4*>(src + r * S + 4 * k);
  }
}

// per-(element, shot) blocks of `blk` values: dst block (e, s) at (e * S2 + s) * blk = src block at
// (e * S + s) * blk, for s < S2
__global__ void k_batch_restride_blocks(int64_t nspec, int S, int S2, int blk, const realw* __restrict__ src,
                                        realw* __restrict__ dst) {
  const int64_t per_e = (int64_t)S2 * blk;
  const int64_t total = nspec * per_e;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / per_e;
    const int64_t w = t - e * per_e;  // s * blk + j
    dst[t] = src[e * (int64_t)S * blk + w];
  }
}

// final displacement of the slots [S2, S) that compaction drops: arch[r * S0 + s] = u[r * S + s]
__global__ void k_batch_archive(int64_t nrow, int S, int S2, int S0, const realw* __restrict__ u,
                                realw* __restrict__ arch) {
  const int nd = S - S2;
  const int64_t total = nrow * nd;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / nd;
    const int s = S2 + (int)(t - r * nd);
    arch[r * S0 + s] = u[r * S + s];
  }
}

static dim3 stride_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > (1 << 20)) b = 1 << 20;
  return dim3((unsigned)b);
}

// ------------------------------------------------------------------------------ host side
template <typename T>
static T* balloc(int64_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)) != hipSuccess)
    fatal("sf_batch", "out of device memory for the batched wavefields");
  return reinterpret_cast<T*>(p);
}
template <typename T>
static void bfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
static T* bupload(const T* h, int64_t n) {
  T* d = balloc<T>(n);
  if (n > 0) SF_CHECK(hipMemcpy(d, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return d;
}

static void drop_graph(Batch* b) {
  // a replay of the executable graph may still run on the stream: wait for it before destroying
  if (b->gexec) {
    SF_CHECK(hipStreamSynchronize(b->m->stream));
    (void)hipGraphExecDestroy(b->gexec);
  }
  b->gexec = nullptr;
  b->graph_steps = 0;
}

// Give the live shots (a prefix of the slots, which are sorted by decreasing record length) rows
// of their own width once at least 8 slots have finished: with S = 32 and 8 live shots, every row
// read would fetch four times the bytes it uses. Slots [S2, S) are finished; their final
// displacement is kept for sf_batch_displ. Kernel arguments change, so captured graphs are dropped.
static void compact(Batch* b, int S2) {
  Mesh* m = b->m;
  hipStream_t st = m->stream;
  drop_graph(b);
  const int64_t nrow = 3 * (int64_t)m->nglob;
  if (!b->u_last) b->u_last = balloc<realw>(nrow * b->S0);
  hipLaunchKernelGGL(k_batch_archive, stride_grid(nrow * (b->S - S2)), dim3(256), 0, st, nrow, b->S, S2, b->S0,
                     b->u, b->u_last);
  SF_LAUNCH_CHECK("k_batch_archive");
  for (realw** f : {&b->u, &b->v, &b->a}) {
    realw* d = balloc<realw>(nrow * S2);
    hipLaunchKernelGGL(k_batch_restride, stride_grid(nrow * (S2 / 4)), dim3(256), 0, st, nrow, b->S, S2, *f, d);
    SF_LAUNCH_CHECK("k_batch_restride");
    SF_CHECK(hipStreamSynchronize(st));
    bfree(*f);
    *f = d;
  }
  if (b->ac) {
    const int64_t na = b->ac_n;
    if (!b->p_last) b->p_last = balloc<realw>(na * b->S0);
    hipLaunchKernelGGL(k_batch_archive, stride_grid(na * (b->S - S2)), dim3(256), 0, st, na, b->S, S2, b->S0,
                       b->pot, b->p_last);
    SF_LAUNCH_CHECK("k_batch_archive");
    for (realw** f : {&b->pot, &b->potd, &b->potdd}) {
      realw* d = balloc<realw>(na * S2);
      hipLaunchKernelGGL(k_batch_restride, stride_grid(na * (S2 / 4)), dim3(256), 0, st, na, b->S, S2, *f, d);
      SF_LAUNCH_CHECK("k_batch_restride");
      SF_CHECK(hipStreamSynchronize(st));
      bfree(*f);
      *f = d;
    }
  }
  if (b->att) {
    const int64_t ne = m->nspec;
    for (auto fb : {std::make_pair(&b->R, 18 * 125), std::make_pair(&b->eps, 6 * 125)}) {
      realw* d = balloc<realw>(ne * S2 * fb.second);
      hipLaunchKernelGGL(k_batch_restride_blocks, stride_grid(ne * S2 * fb.second), dim3(256), 0, st, ne, b->S, S2,
                         fb.second, *fb.first, d);
      SF_LAUNCH_CHECK("k_batch_restride_blocks");
      SF_CHECK(hipStreamSynchronize(st));
      bfree(*fb.first);
      *fb.first = d;
    }
  }
  b->S = S2;
}

static void maybe_compact(Batch* b) {
  int nlive = 0;
  for (int s = 0; s < b->S; s++) nlive += b->h_nstep[s] >= b->h_it ? 1 : 0;
  const int S2 = std::max(8, (nlive + 7) / 8 * 8);
  if (S2 < b->S) compact(b, S2);
}

// back to rows of the full width (a rewind after compaction); the contents are zeroed next
static void restore_width(Batch* b) {
  if (b->S == b->S0) return;
  Mesh* m = b->m;
  drop_graph(b);
  SF_CHECK(hipStreamSynchronize(m->stream));
  const int64_t n = 3 * (int64_t)m->nglob * b->S0;
  for (realw** f : {&b->u, &b->v, &b->a}) {
    bfree(*f);
    *f = balloc<realw>(n);
  }
  if (b->ac) {
    for (realw** f : {&b->pot, &b->potd, &b->potdd}) {
      bfree(*f);
      *f = balloc<realw>(b->ac_n * b->S0);
    }
  }
  if (b->att) {
    bfree(b->R);
    bfree(b->eps);
    b->R = balloc<realw>(18 * (int64_t)m->nspec * 125 * b->S0);
    b->eps = balloc<realw>(6 * (int64_t)m->nspec * 125 * b->S0);
  }
  b->S = b->S0;
}

static void zero_state(Batch* b) {
  restore_width(b);
  Mesh* m = b->m;
  const int64_t n = 3 * (int64_t)m->nglob * b->S;
  for (realw* p : {b->u, b->v, b->a})
    hipLaunchKernelGGL(k_batch_fill, sf_grid(n, 256), dim3(256), 0, m->stream, n, p);
  if (b->ac) {
    const int64_t na = b->ac_n * b->S;
    for (realw* p : {b->pot, b->potd, b->potdd})
      hipLaunchKernelGGL(k_batch_fill, sf_grid(na, 256), dim3(256), 0, m->stream, na, p);
  }
  if (b->att) {
    const int64_t nr = 18 * (int64_t)m->nspec * 125 * b->S, ne = 6 * (int64_t)m->nspec * 125 * b->S;
    hipLaunchKernelGGL(k_batch_fill, sf_grid(nr, 256), dim3(256), 0, m->stream, nr, b->R);
    hipLaunchKernelGGL(k_batch_fill, sf_grid(ne, 256), dim3(256), 0, m->stream, ne, b->eps);
  }
  if (b->seis) {
    const int64_t ns = (int64_t)b->max_nstep * b->nrec * 3;
    hipLaunchKernelGGL(k_batch_fill, sf_grid(ns, 256), dim3(256), 0, m->stream, ns, b->seis);
  }
  SF_LAUNCH_CHECK("k_batch_fill");
  b->h_it = 1;
  SF_CHECK(hipMemcpyAsync(b->d_it, &b->h_it, sizeof(int), hipMemcpyHostToDevice, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
}

// Point flags of the fused mode: for each global point, the element of its lowest colour writes
// first (unless Stacey or a source adds to the point before the colours), the element of its
// highest colour finishes the Newmark update and clears `a` where those passes add.
static int point_flags(int nglob, int nelem, const int* ibool0, const int* elems, const int* color,
                       const unsigned char* forced, unsigned char* flags) {
  std::vector<int> cmin((size_t)nglob, 1 << 30), cmax((size_t)nglob, -1), nlast((size_t)nglob, 0);
  for (int q = 0; q < nelem; q++)
    for (int p = 0; p < 125; p++) {
      const int g = ibool0[(size_t)elems[q] * 125 + p];
      cmin[g] = std::min(cmin[g], color[q]);
      cmax[g] = std::max(cmax[g], color[q]);
    }
  for (int q = 0; q < nelem; q++) {
    const int e = elems[q], c = color[q];
    for (int p = 0; p < 125; p++) {
      const int g = ibool0[(size_t)e * 125 + p];
      unsigned char f = forced[g] ? BF_ZERO : 0;
      if (c == cmin[g] && !forced[g]) f |= BF_FIRST;
      if (c == cmax[g]) {
        f |= BF_LAST;
        nlast[g]++;
      }
      flags[(size_t)e * 125 + p] = f;
    }
  }
  for (int g = 0; g < nglob; g++)
    if (cmax[g] >= 0 && nlast[g] != 1) return -1;  // two elements of one colour share the point
  return 0;
}

static void build_point_flags(Batch* b) {
  Mesh* m = b->m;
  ensure_element_colors(m);
  const int ne = m->nspec_elastic;
  std::vector<int> lst(ne), hib((size_t)m->nspec * 125), col(ne);
  SF_CHECK(hipMemcpy(lst.data(), m->el_color_elems, sizeof(int) * ne, hipMemcpyDeviceToHost));
  SF_CHECK(hipMemcpy(hib.data(), m->ibool, sizeof(int) * hib.size(), hipMemcpyDeviceToHost));
  for (int c = 0; c < m->el_ncolor; c++)
    for (int q = m->el_color_off[c]; q < m->el_color_off[c + 1]; q++) col[q] = c;
  std::vector<unsigned char> forced((size_t)m->nglob, 0);
  if (m->stacey && m->abs_el.n > 0) {
    std::vector<int> ig(m->abs_el.n);
    SF_CHECK(hipMemcpy(ig.data(), m->abs_el.iglob, sizeof(int) * ig.size(), hipMemcpyDeviceToHost));
    for (int g : ig) forced[g] = 1;
  }
  if (b->ac && m->coupling.n > 0) {
    // the solid side of the coupling adds to these points before the colours
    std::vector<int> ig(m->coupling.n);
    SF_CHECK(hipMemcpy(ig.data(), m->coupling.iglob, sizeof(int) * ig.size(), hipMemcpyDeviceToHost));
    for (int g : ig) forced[g] = 1;
  }
  for (const Batch::Src& q : b->srcs)
    for (int p = 0; p < 125; p++) forced[hib[(size_t)q.ispec * 125 + p]] = 1;
  std::vector<unsigned char> fl((size_t)m->nspec * 125, 0);
  if (point_flags(m->nglob, ne, hib.data(), lst.data(), col.data(), forced.data(), fl.data()) != 0)
    fatal("sf_batch", "element colouring shares a point inside a colour");
  bfree(b->flags);
  b->flags = bupload(fl.data(), (int64_t)fl.size());
  b->flags_version = m->color_version;
}

static void enqueue_stiffness(Batch* b, hipStream_t st) {
  Mesh* m = b->m;
  BatchArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ngroups = b->S / b->sc;
  A.S = b->S;
  A.ibool = m->ibool;
  A.irregular = m->irregular;
  A.xix_regular = m->xix_regular;
  A.jacobian_regular = m->jacobian_regular;
  A.jac = m->jac;
  A.jac_plane = (int64_t)m->nspec_irregular * 125;
  A.kappamu = m->kappamu;
  A.u = b->u;
  A.a = b->a;
  A.nstep = b->d_nstep;
  A.it = b->d_it;
  A.v = b->v;
  A.rmass = m->rmass;
  A.flags = b->flags;
  A.dt = b->dt;
  A.dt2 = b->dt2;
  A.dtsq2 = b->dtsq2;
  if (b->att) {
    A.R = b->R;
    A.eps = b->eps;
    A.fc_index = m->fc_index;
    A.fc_table = m->fc_table;
    A.fc_rows = m->fc_rows;
    A.fc = m->factor_common;
    A.fck = m->factor_common_kappa;
    A.nloc = (int64_t)m->nspec * 125;
    for (int s = 0; s < 3; s++) {
      A.alpha[s] = m->alphaval[s];
      A.beta[s] = m->betaval[s];
      A.gamma[s] = m->gammaval[s];
    }
    for (int c = 0; c < m->el_ncolor; c++) {
      A.elems = m->el_color_elems + m->el_color_off[c];
      A.nelem = m->el_color_off[c + 1] - m->el_color_off[c];
      launch_visco_batch(st, A, b->sc, b->mode, m->exact != 0);
    }
  } else if (b->mode == 2 || b->mode == 3) {
    for (int c = 0; c < m->el_ncolor; c++) {
      A.elems = m->el_color_elems + m->el_color_off[c];
      A.nelem = m->el_color_off[c + 1] - m->el_color_off[c];
      launch_elastic_batch(st, A, b->sc, b->mode, b->overlay, m->exact != 0);
    }
  } else {
    A.elems = m->elastic_elems;
    A.nelem = m->nspec_elastic;
    launch_elastic_batch(st, A, b->sc, 0, b->overlay, m->exact != 0);
  }
}

// one time step; `first` selects the plain predictor (no pending corrector)
static void enqueue_seismo(Batch* b, hipStream_t st) {
  if (b->nrec <= 0) return;
  hipLaunchKernelGGL(k_batch_seismo, dim3(b->nrec), dim3(64), 0, st, b->nrec, b->S, b->rec_slot, b->rec_ig,
                     b->rec_h, b->rec_nu, b->d_nstep, b->d_it, b->u, b->seis);
  SF_LAUNCH_CHECK("k_batch_seismo");
}

// launches over (face point of one round, live shot) pairs
static dim3 pair_grid(int cnt, int S, const char* what) {
  if ((int64_t)cnt * S >= 0x7fffffff) fatal("sf_batch", what);
  return sf_grid((int64_t)cnt * S, 256);
}

// the fluid part of a step (the acoustic calls of iterate_time.F90 before the elastic ones):
// predictor, free surface, Stacey, coupling (fluid side), stiffness (all elements are inner at
// NPROC = 1, so the stiffness runs in phase 2, after the boundary terms of phase 1), corrector,
// free surface
static void enqueue_fluid(Batch* b, hipStream_t st) {
  Mesh* m = b->m;
  const int S = b->S;
  const int rpb = 256 / (S / 4);
  const int64_t na = b->ac_n;
  const dim3 rows = sf_grid((na + rpb - 1) / rpb, 1);
  hipLaunchKernelGGL(k_batch_ac_newmark, rows, dim3(256), 0, st, na, S, rpb, b->d_nstep, b->d_it, b->pot, b->potd,
                     b->potdd, m->rmass_ac + b->ac_g0, b->dt, b->dt2, b->dtsq2, 0);
  SF_LAUNCH_CHECK("k_batch_ac_newmark");
  const FacePoints& fs = m->free_ac;
  if (fs.n > 0)
    hipLaunchKernelGGL(k_batch_free_surface, pair_grid(fs.n, S, "too many free-surface points x shots"), dim3(256),
                       0, st, fs.n, S, b->ac_g0, fs.iglob, b->d_nstep, b->d_it, b->pot, b->potd, b->potdd);
  if (m->stacey) {
    const FacePoints& f = m->abs_ac;
    for (int r = 0; r < f.nrounds; r++) {
      const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
      if (cnt <= 0) continue;
      hipLaunchKernelGGL(k_batch_stacey_ac, pair_grid(cnt, S, "too many fluid Stacey points x shots"), dim3(256), 0,
                         st, cnt, S, b->ac_g0, f.iglob + o, f.jacw + o, f.imp_p + o, f.imp_s + o, b->d_nstep,
                         b->d_it, b->potd, b->potdd);
    }
    SF_LAUNCH_CHECK("k_batch_stacey_ac");
  }
  const FacePoints& cf = m->coupling;
  for (int r = 0; r < cf.nrounds; r++) {
    const int o = cf.round_off[r], cnt = cf.round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_batch_coupling_ac_el, pair_grid(cnt, S, "too many coupling points x shots"), dim3(256), 0,
                       st, cnt, S, b->ac_g0, cf.iglob + o, cf.normal + 3 * o, cf.jacw + o, b->d_nstep, b->d_it, b->u,
                       b->potdd);
  }
  SF_LAUNCH_CHECK("k_batch_coupling_ac_el");
  BatchAcArgs A;
  std::memset(&A, 0, sizeof(A));
  A.ngroups = S / b->sc;
  A.S = S;
  A.g0 = b->ac_g0;
  A.ibool = m->ibool;
  A.irregular = m->irregular;
  A.xix_regular = m->xix_regular;
  A.jacobian_regular = m->jacobian_regular;
  A.jac = m->jac;
  A.jac_plane = (int64_t)m->nspec_irregular * 125;
  A.rho = m->rho;
  A.pot = b->pot;
  A.potdd = b->potdd;
  A.nstep = b->d_nstep;
  A.it = b->d_it;
  if (b->mode == 0) {
    A.elems = m->acoustic_elems;
    A.nelem = m->nspec_acoustic;
    launch_acoustic_batch(st, A, b->sc, 0, m->exact != 0);
  } else {
    for (int c = 0; c < m->ac_ncolor; c++) {
      A.elems = m->ac_color_elems + m->ac_color_off[c];
      A.nelem = m->ac_color_off[c + 1] - m->ac_color_off[c];
      launch_acoustic_batch(st, A, b->sc, 2, m->exact != 0);
    }
  }
  hipLaunchKernelGGL(k_batch_ac_newmark, rows, dim3(256), 0, st, na, S, rpb, b->d_nstep, b->d_it, b->pot, b->potd,
                     b->potdd, m->rmass_ac + b->ac_g0, b->dt, b->dt2, b->dtsq2, 1);
  SF_LAUNCH_CHECK("k_batch_ac_newmark");
  if (fs.n > 0)
    hipLaunchKernelGGL(k_batch_free_surface, pair_grid(fs.n, S, "too many free-surface points x shots"), dim3(256),
                       0, st, fs.n, S, b->ac_g0, fs.iglob, b->d_nstep, b->d_it, b->pot, b->potd, b->potdd);
  SF_LAUNCH_CHECK("k_batch_free_surface");
}

static void enqueue_step(Batch* b, hipStream_t st, bool first) {
  Mesh* m = b->m;
  const bool fused = b->mode == 3;
  if (fused) {
    enqueue_seismo(b, st);  // u already holds this step's predicted displacement
  } else {
    const int64_t nrow = 3 * (int64_t)m->nglob;
    const realw dt = b->dt, dt2 = b->dt2, dtsq2 = b->dtsq2;
    const int rpb = 256 / (b->S / 4);
    hipLaunchKernelGGL(k_batch_update, sf_grid((nrow + rpb - 1) / rpb, 1), dim3(256), 0, st, nrow, b->S, rpb,
                       b->d_nstep, b->d_it, b->u, b->v, b->a, m->rmass, dt2, dt, dt2, dtsq2, first ? 1 : 0);
    SF_LAUNCH_CHECK("k_batch_update");
  }
  if (b->ac) enqueue_fluid(b, st);
  if (m->stacey) {
    const FacePoints& f = m->abs_el;
    for (int r = 0; r < f.nrounds; r++) {
      const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
      if (cnt <= 0) continue;
      if ((int64_t)cnt * b->S >= 0x7fffffff) fatal("sf_batch", "too many Stacey points x shots");
      hipLaunchKernelGGL(k_batch_stacey, sf_grid((int64_t)cnt * b->S, 256), dim3(256), 0, st, cnt, b->S,
                         f.iglob + o, f.normal + 3 * o, f.jacw + o, f.imp_p + o, f.imp_s + o, b->d_nstep, b->d_it,
                         b->v, b->a);
    }
    SF_LAUNCH_CHECK("k_batch_stacey");
  }
  if (b->ac) {
    const FacePoints& cf = m->coupling;
    for (int r = 0; r < cf.nrounds; r++) {
      const int o = cf.round_off[r], cnt = cf.round_off[r + 1] - o;
      if (cnt <= 0) continue;
      hipLaunchKernelGGL(k_batch_coupling_el_ac, pair_grid(cnt, b->S, "too many coupling points x shots"), dim3(256),
                         0, st, cnt, b->S, b->ac_g0, cf.iglob + o, cf.normal + 3 * o, cf.jacw + o, b->d_nstep,
                         b->d_it, b->potdd, b->a);
    }
    SF_LAUNCH_CHECK("k_batch_coupling_el_ac");
  }
  for (int r = 0; r < b->src_nrounds; r++) {
    const int o = b->src_round_off[r], cnt = b->src_round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_batch_sources, sf_grid(cnt, 128), dim3(128), 0, st, cnt, b->S, b->max_nstep, b->src_ig + o,
                       b->src_slot + o, b->src_id + o, b->src_ploc + o, b->src_arr, b->src_stf, b->d_nstep, b->d_it,
                       b->a);
    SF_LAUNCH_CHECK("k_batch_sources");
  }
  const bool timed = b->timing && b->ev_used + 2 <= (int)b->ev.size();
  if (timed) SF_CHECK(hipEventRecord(b->ev[b->ev_used], st));
  enqueue_stiffness(b, st);
  if (timed) {
    SF_CHECK(hipEventRecord(b->ev[b->ev_used + 1], st));
    b->ev_used += 2;
  }
  if (!fused) enqueue_seismo(b, st);
  hipLaunchKernelGGL(k_batch_tick, dim3(1), dim3(64), 0, st, b->d_it);
  SF_LAUNCH_CHECK("k_batch_tick");
}

// advance n steps (eagerly for the first step and for a tail shorter than the captured group)
static void run_steps(Batch* b, int n, int graph_steps) {
  Mesh* m = b->m;
  n = std::min(n, b->max_nstep - b->h_it + 1);
  if (b->mode == 3 && (!b->flags || b->flags_version != m->color_version)) {
    drop_graph(b);
    build_point_flags(b);
  }
  while (n > 0) {
    maybe_compact(b);
    if (b->h_it == 1 || graph_steps <= 1 || n < graph_steps) {
      enqueue_step(b, m->stream, b->h_it == 1);
      b->h_it++;
      n--;
      continue;
    }
    if (!b->gexec || b->graph_steps != graph_steps) {
      drop_graph(b);
      hipGraph_t g;
      SF_CHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < graph_steps; k++) enqueue_step(b, m->stream, false);
      SF_CHECK(hipStreamEndCapture(m->stream, &g));
      SF_CHECK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
      SF_CHECK(hipGraphDestroy(g));
      b->graph_steps = graph_steps;
    }
    SF_CHECK(hipGraphLaunch(b->gexec, m->stream));
    b->h_it += graph_steps;
    n -= graph_steps;
  }
}

}  // namespace sf

using namespace sf;

// Create a batch of `nshots` shots with record lengths nstep[i] on a prepared elastic mesh (after
// prepare_constants_device_ and prepare_fields_elastic_device_; sources/receivers of the mesh
// itself are not used), time step constants as the host computes them (prepare_timerun.F90:
// deltat, deltatover2 = deltat/2, deltatsqover2 = deltat^2/2). Returns the handle; shots are
// numbered 0..nshots-1 by the caller.
extern "C" long sf_batch_create(long* Mesh_pointer, int nshots, const int* nstep, realw deltat, realw deltatover2,
                                realw deltatsqover2) {
  Mesh* m = from_handle(Mesh_pointer);
  if (nshots <= 0) fatal("sf_batch_create", "need at least one shot");
  if (!m->rmass || !m->kappamu || m->nspec_elastic <= 0) fatal("sf_batch_create", "mesh has no prepared elastic domain");
  if (m->pml_n > 0) fatal("sf_batch_create", "C-PML meshes are not supported by the batched engine");
  if (m->nspec_acoustic > 0 && (!m->ac[0].pot || !m->rmass_ac))
    fatal("sf_batch_create", "mesh has fluid elements but prepare_fields_acoustic_device was not called");
  if (m->attenuation && !m->fc_index && !m->factor_common)
    fatal("sf_batch_create", "attenuation factors not prepared");
  Batch* b = new Batch();
  b->m = m;
  b->dt = deltat;
  b->dt2 = deltatover2;
  b->dtsq2 = deltatsqover2;
  b->nshots = nshots;
  ensure_element_colors(m);
  if (m->nspec_acoustic > 0) {
    ensure_acoustic_colors(m);
    b->ac = true;
    b->ac_g0 = m->ac_g0;
    b->ac_n = m->ac_n;
  }
  b->S = (nshots + 7) / 8 * 8;
  if (b->S > 1024) fatal("sf_batch_create", "at most 1024 shots per batch");
  b->nstep_user.assign(nstep, nstep + nshots);
  for (int i = 0; i < nshots; i++)
    if (nstep[i] <= 0) fatal("sf_batch_create", "record lengths must be positive");
  b->max_nstep = *std::max_element(nstep, nstep + nshots);
  std::vector<int> order(nshots);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nstep[x] > nstep[y]; });
  b->slot_of.assign(nshots, 0);
  std::vector<int> ns(b->S, 0);
  for (int k = 0; k < nshots; k++) {
    b->slot_of[order[k]] = k;
    ns[k] = nstep[order[k]];
  }
  b->d_nstep = bupload(ns.data(), b->S);
  b->h_nstep = ns;
  b->S0 = b->S;
  b->d_it = balloc<int>(1);
  const int64_t n = 3 * (int64_t)m->nglob * b->S;
  b->u = balloc<realw>(n);
  b->v = balloc<realw>(n);
  b->a = balloc<realw>(n);
  if (b->ac) {
    b->pot = balloc<realw>(b->ac_n * b->S);
    b->potd = balloc<realw>(b->ac_n * b->S);
    b->potdd = balloc<realw>(b->ac_n * b->S);
  }
  if (m->attenuation) {
    b->att = true;
    b->mode = 3;
    b->R = balloc<realw>(18 * (int64_t)m->nspec * 125 * b->S);
    b->eps = balloc<realw>(6 * (int64_t)m->nspec * 125 * b->S);
  }
  return reinterpret_cast<long>(b);
}

// Add a moment-tensor or force source to a shot: element ispec (0-based), sourcearray in the
// reference's (NDIM,NGLLX,NGLLY,NGLLZ) order (component fastest), and the source-time-function
// value used at every step of the shot (stf[it-1] = real(stf_pre_compute) of step it).
extern "C" void sf_batch_add_source(long h, int shot, int ispec, const realw* sourcearray, const realw* stf) {
  Batch* b = batch_of(h);
  if (shot < 0 || shot >= b->nshots) fatal("sf_batch_add_source", "shot index out of range");
  if (ispec < 0 || ispec >= b->m->nspec) fatal("sf_batch_add_source", "element index out of range");
  if (!element_is_elastic(b->m, ispec))
    fatal("sf_batch_add_source", "batched sources must lie in elastic elements (single-shot runs take fluid ones)");
  Batch::Src s;
  s.shot = shot;
  s.ispec = ispec;
  s.arr.assign(sourcearray, sourcearray + 375);
  s.stf.assign(stf, stf + b->nstep_user[shot]);
  b->srcs.push_back(std::move(s));
  b->ready = false;
}

// Add a receiver to a shot: element ispec (0-based), Lagrange interpolants hxi/heta/hgamma (5 each,
// double) and the rotation nu (3x3, nu_rec(3,3) column-major as the reference passes it).
extern "C" void sf_batch_add_receiver(long h, int shot, int ispec, const double* hxi, const double* heta,
                                      const double* hgamma, const double* nu) {
  Batch* b = batch_of(h);
  if (shot < 0 || shot >= b->nshots) fatal("sf_batch_add_receiver", "shot index out of range");
  if (ispec < 0 || ispec >= b->m->nspec) fatal("sf_batch_add_receiver", "element index out of range");
  if (!element_is_elastic(b->m, ispec))
    fatal("sf_batch_add_receiver",
          "batched receivers must lie in elastic elements (single-shot runs record fluid ones)");
  Batch::Rec r;
  r.shot = shot;
  r.ispec = ispec;
  for (int p = 0; p < 125; p++) r.h[p] = hxi[p % 5] * heta[(p / 5) % 5] * hgamma[p / 25];
  // rotated component a = sum_b nu(a, b) u_b; nu_rec(3,3) column-major: nu(a,b) at a + 3 b
  for (int x = 0; x < 3; x++)
    for (int y = 0; y < 3; y++) r.nu[3 * x + y] = nu[x + 3 * y];
  b->recs.push_back(r);
  b->ready = false;
}

// Upload sources and receivers, zero the wavefields and seismograms, set the step counter to 1.
extern "C" void sf_batch_finalize(long h) {
  Batch* b = batch_of(h);
  Mesh* m = b->m;
  SF_CHECK(hipStreamSynchronize(m->stream));
  drop_graph(b);
  std::vector<int> hib((size_t)m->nspec * 125);
  SF_CHECK(hipMemcpy(hib.data(), m->ibool, sizeof(int) * hib.size(), hipMemcpyDeviceToHost));
  // sources: round of an entry = index of its source within its shot
  {
    std::vector<int> k_in_shot(b->srcs.size()), count(b->nshots, 0);
    for (size_t q = 0; q < b->srcs.size(); q++) k_in_shot[q] = count[b->srcs[q].shot]++;
    const int nr = b->srcs.empty() ? 0 : *std::max_element(count.begin(), count.end());
    std::vector<int> ig, slot, sid, ploc;
    b->src_round_off.assign(nr + 1, 0);
    for (int r = 0; r < nr; r++) {
      b->src_round_off[r] = (int)ig.size();
      for (size_t q = 0; q < b->srcs.size(); q++) {
        if (k_in_shot[q] != r) continue;
        for (int p = 0; p < 125; p++) {
          ig.push_back(hib[(size_t)b->srcs[q].ispec * 125 + p]);
          slot.push_back(b->slot_of[b->srcs[q].shot]);
          sid.push_back((int)q);
          ploc.push_back(p);
        }
      }
    }
    b->src_round_off[nr] = (int)ig.size();
    b->src_nrounds = nr;
    b->src_n = (int)ig.size();
    std::vector<realw> arr(b->srcs.size() * 375), stf(b->srcs.size() * (size_t)b->max_nstep, 0.f);
    for (size_t q = 0; q < b->srcs.size(); q++) {
      // (c, i, j, k) with c fastest -> [p][c]
      for (int p = 0; p < 125; p++)
        for (int c = 0; c < 3; c++) arr[q * 375 + p * 3 + c] = b->srcs[q].arr[c + 3 * p];
      std::copy(b->srcs[q].stf.begin(), b->srcs[q].stf.end(), stf.begin() + q * b->max_nstep);
    }
    bfree(b->src_ig); bfree(b->src_slot); bfree(b->src_id); bfree(b->src_ploc); bfree(b->src_arr); bfree(b->src_stf);
    b->src_ig = bupload(ig.data(), (int64_t)ig.size());
    b->src_slot = bupload(slot.data(), (int64_t)slot.size());
    b->src_id = bupload(sid.data(), (int64_t)sid.size());
    b->src_ploc = bupload(ploc.data(), (int64_t)ploc.size());
    b->src_arr = bupload(arr.data(), (int64_t)arr.size());
    b->src_stf = bupload(stf.data(), (int64_t)stf.size());
  }
  // receivers, device order = order of addition
  {
    b->nrec = (int)b->recs.size();
    std::vector<int> slot(b->nrec), ig((size_t)b->nrec * 125);
    std::vector<double> hh((size_t)b->nrec * 125), nu((size_t)b->nrec * 9);
    for (int r = 0; r < b->nrec; r++) {
      const Batch::Rec& R = b->recs[r];
      slot[r] = b->slot_of[R.shot];
      for (int p = 0; p < 125; p++) {
        ig[(size_t)r * 125 + p] = hib[(size_t)R.ispec * 125 + p];
        hh[(size_t)r * 125 + p] = R.h[p];
      }
      for (int q = 0; q < 9; q++) nu[(size_t)r * 9 + q] = R.nu[q];
    }
    bfree(b->rec_slot); bfree(b->rec_ig); bfree(b->rec_h); bfree(b->rec_nu); bfree(b->seis);
    b->rec_slot = bupload(slot.data(), (int64_t)slot.size());
    b->rec_ig = bupload(ig.data(), (int64_t)ig.size());
    b->rec_h = bupload(hh.data(), (int64_t)hh.size());
    b->rec_nu = bupload(nu.data(), (int64_t)nu.size());
    b->seis = balloc<realw>((int64_t)b->max_nstep * std::max(1, b->nrec) * 3);
  }
  b->flags_version = -1;  // sources changed: the fused mode's cleared points change
  b->ready = true;
  zero_state(b);
}

// shots per workgroup (4 or 8; divides the padded shot count) and assembly (0 atomic, 2 coloured,
// 3 coloured with the Newmark update fused). The modes keep different intermediate states, so the
// batch must be at step 1 (new, or after sf_batch_reset).
extern "C" void sf_batch_set_options(long h, int shots_per_group, int assembly) {
  Batch* b = batch_of(h);
  if (shots_per_group != 4 && shots_per_group != 8)
    fatal("sf_batch_set_options", "shots per workgroup must be 4 or 8");
  if (assembly != 0 && assembly != 2 && assembly != 3)
    fatal("sf_batch_set_options", "assembly must be 0 (atomic), 2 (coloured) or 3 (coloured, fused update)");
  if (b->h_it != 1) fatal("sf_batch_set_options", "options change only at step 1 (call sf_batch_reset first)");
  if (b->att && assembly == 0) fatal("sf_batch_set_options", "viscoelastic batches use coloured assembly (2 or 3)");
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  drop_graph(b);
  b->sc = shots_per_group;
  b->mode = assembly;
}

// Kernel variants that give the same results: "lds_overlay" 0/1 (the displacement tile shares LDS
// with the stress products). Unknown names are fatal.
extern "C" void sf_batch_set_tuning(long h, const char* name, int value) {
  Batch* b = batch_of(h);
  if (std::strcmp(name, "lds_overlay") == 0) {
    if (value < 0 || value > 1) fatal("sf_batch_set_tuning", "lds_overlay must be 0 or 1");
    SF_CHECK(hipStreamSynchronize(b->m->stream));
    drop_graph(b);
    b->overlay = value;
  } else {
    fatal("sf_batch_set_tuning", "unknown tuning switch");
  }
}

extern "C" int sf_batch_get_tuning(long h, const char* name) {
  Batch* b = batch_of(h);
  if (std::strcmp(name, "lds_overlay") == 0) return b->overlay;
  fatal("sf_batch_get_tuning", "unknown tuning switch");
  return 0;
}

extern "C" void sf_batch_get_options(long h, int* shots_per_group, int* assembly) {
  Batch* b = batch_of(h);
  *shots_per_group = b->sc;
  *assembly = b->mode;
}

// advance `nsteps` steps (at most to the longest record); `graph_steps` > 1 captures that many
// steps into one hipGraph and replays it
extern "C" void sf_batch_run(long h, int nsteps, int graph_steps) {
  Batch* b = batch_of(h);
  if (!b->ready) fatal("sf_batch_run", "sf_batch_finalize not called after adding sources/receivers");
  run_steps(b, nsteps, graph_steps);
}

extern "C" int sf_batch_current_step(long h) { return batch_of(h)->h_it; }

// Seismograms of one shot: out[(it-1) * nrec_shot * 3 + r * 3 + c] for its receivers in the order
// they were added, for it = 1..nstep of the shot. Synchronises the stream.
extern "C" void sf_batch_seismograms(long h, int shot, realw* out) {
  Batch* b = batch_of(h);
  if (shot < 0 || shot >= b->nshots) fatal("sf_batch_seismograms", "shot index out of range");
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  std::vector<int> mine;
  for (int r = 0; r < b->nrec; r++)
    if (b->recs[r].shot == shot) mine.push_back(r);
  if (mine.empty()) return;
  const int ns = b->nstep_user[shot];
  std::vector<realw> all((size_t)ns * b->nrec * 3);
  SF_CHECK(hipMemcpy(all.data(), b->seis, sizeof(realw) * all.size(), hipMemcpyDeviceToHost));
  const size_t nr = mine.size();
  for (int t = 0; t < ns; t++)
    for (size_t q = 0; q < nr; q++)
      for (int c = 0; c < 3; c++) out[((size_t)t * nr + q) * 3 + c] = all[((size_t)t * b->nrec + mine[q]) * 3 + c];
}

// Displacement of one shot (3*NGLOB, xyz interleaved) copied to the host; synchronises.
extern "C" void sf_batch_displ(long h, int shot, realw* out) {
  Batch* b = batch_of(h);
  if (shot < 0 || shot >= b->nshots) fatal("sf_batch_displ", "shot index out of range");
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  const int64_t n = 3 * (int64_t)b->m->nglob;
  const int s = b->slot_of[shot];
  const bool dropped = s >= b->S;  // finished and compacted away: its final state was archived
  const int stride = dropped ? b->S0 : b->S;
  std::vector<realw> all((size_t)n * stride);
  SF_CHECK(hipMemcpy(all.data(), dropped ? b->u_last : b->u, sizeof(realw) * all.size(), hipMemcpyDeviceToHost));
  for (int64_t q = 0; q < n; q++) out[q] = all[(size_t)q * stride + s];
}

// Potential of one shot over all global points (zero outside the acoustic range; all zero on a mesh
// without fluid elements). Synchronises the stream.
extern "C" void sf_batch_potential(long h, int shot, realw* out) {
  Batch* b = batch_of(h);
  if (shot < 0 || shot >= b->nshots) fatal("sf_batch_potential", "shot index out of range");
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  const int64_t n = b->m->nglob;
  for (int64_t q = 0; q < n; q++) out[q] = 0.f;
  if (!b->ac) return;
  const int s = b->slot_of[shot];
  const bool dropped = s >= b->S;
  const int stride = dropped ? b->S0 : b->S;
  std::vector<realw> all((size_t)b->ac_n * stride);
  SF_CHECK(hipMemcpy(all.data(), dropped ? b->p_last : b->pot, sizeof(realw) * all.size(), hipMemcpyDeviceToHost));
  for (int64_t q = 0; q < b->ac_n; q++) out[b->ac_g0 + q] = all[(size_t)q * stride + s];
}

// Rewind to step 1 with zero wavefields and seismograms (keeps sources, receivers and options).
extern "C" void sf_batch_reset(long h) {
  Batch* b = batch_of(h);
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  zero_state(b);
}

// Time `nsteps` eager steps from the current state: average milliseconds of the stiffness launches
// per step (HIP events around them on the mesh stream) and of the whole step.
extern "C" void sf_batch_time(long h, int nsteps, double* stiffness_ms, double* step_ms) {
  Batch* b = batch_of(h);
  Mesh* m = b->m;
  nsteps = std::min(nsteps, b->max_nstep - b->h_it + 1);
  if (nsteps <= 0) fatal("sf_batch_time", "no steps left to time");
  b->ev.resize(2 * (size_t)nsteps + 2);
  for (auto& e : b->ev) SF_CHECK(hipEventCreate(&e));
  b->ev_used = 0;
  b->timing = true;
  hipEvent_t t0 = b->ev[2 * nsteps], t1 = b->ev[2 * nsteps + 1];
  b->ev.resize(2 * (size_t)nsteps);
  SF_CHECK(hipEventRecord(t0, m->stream));
  run_steps(b, nsteps, 1);
  SF_CHECK(hipEventRecord(t1, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
  b->timing = false;
  double tot = 0.0;
  for (int k = 0; k + 1 < b->ev_used; k += 2) {
    float ms = 0.f;
    SF_CHECK(hipEventElapsedTime(&ms, b->ev[k], b->ev[k + 1]));
    tot += ms;
  }
  float all = 0.f;
  SF_CHECK(hipEventElapsedTime(&all, t0, t1));
  *stiffness_ms = tot / std::max(1, b->ev_used / 2);
  *step_ms = (double)all / nsteps;
  for (auto& e : b->ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->ev.clear();
  b->ev_used = 0;
}

// Autotuner: times every (assembly, shots per workgroup, LDS overlay) candidate for `trial_steps`
// eager steps from rest (after a warm-up; a candidate whose first step takes over twice the best
// so far is dropped), keeps the fastest and rewinds the batch. Returns the chosen step time in ms.
extern "C" double sf_batch_autotune(long h, int trial_steps) {
  Batch* b = batch_of(h);
  if (!b->ready) fatal("sf_batch_autotune", "sf_batch_finalize first");
  double best = 1e300;
  int best_sc = b->sc, best_mode = b->mode, best_ov = b->overlay;
  for (int mode : {3, 2, 0})
    for (int sc : {4, 8})
    for (int ov : {1, 0}) {
      if (b->S % sc) continue;
      if (b->att && (mode == 0 || ov != 1)) continue;  // coloured only; no overlay switch
      b->sc = sc;
      b->mode = mode;
      b->overlay = ov;
      drop_graph(b);
      zero_state(b);
      run_steps(b, 2, 1);  // warm-up (first step runs the plain predictor)
      double ks = 0.0, st = 0.0;
      sf_batch_time(h, 1, &ks, &st);
      if (st > 2.0 * best) continue;  // far behind the best so far: not worth the full trial
      sf_batch_time(h, trial_steps, &ks, &st);
      if (st < best) {
        best = st;
        best_sc = sc;
        best_mode = mode;
        best_ov = ov;
      }
    }
  b->sc = best_sc;
  b->mode = best_mode;
  b->overlay = best_ov;
  drop_graph(b);
  zero_state(b);
  return best;
}

extern "C" void sf_batch_destroy(long h) {
  Batch* b = batch_of(h);
  SF_CHECK(hipStreamSynchronize(b->m->stream));
  drop_graph(b);
  bfree(b->d_nstep); bfree(b->d_it); bfree(b->u); bfree(b->v); bfree(b->a); bfree(b->R); bfree(b->eps);
  bfree(b->u_last);
  bfree(b->pot); bfree(b->potd); bfree(b->potdd); bfree(b->p_last);
  bfree(b->flags);
  bfree(b->src_ig); bfree(b->src_slot); bfree(b->src_id); bfree(b->src_ploc); bfree(b->src_arr); bfree(b->src_stf);
  bfree(b->rec_slot); bfree(b->rec_ig); bfree(b->rec_h); bfree(b->rec_nu); bfree(b->seis);
  delete b;
}

// Host-only (no device needed): the fused mode's per-local-point flags for elements elems[q] of
// colour color[q] (colours run in increasing order) and the points that Stacey or sources add to
// before the colours (forced[g] != 0): bit 1 first writer, bit 2 finisher, bit 4 cleared by the
// finisher. flags has 125 entries per element of ibool0. Returns 0, or -1 if two elements of one
// colour share a point.
extern "C" int sf_batch_point_flags(int nglob, int nelem, const int* ibool0, const int* elems, const int* color,
                                    const unsigned char* forced, unsigned char* flags) {
  return point_flags(nglob, nelem, ibool0, elems, color, forced, flags);
}
