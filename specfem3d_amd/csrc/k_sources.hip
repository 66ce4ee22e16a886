// k_sources.hip -- point sources, adjoint sources and receiver interpolation.
//
// Sources and adjoint sources are flattened to (owner, point) pairs sorted into rounds like
// the face points (FacePoints.ispec holds the owner: source or adjoint-receiver index), so
// points shared by two sources are summed in the reference's source order without atomics.
// Built with -ffp-contract=off (reference rounding).
#include "kernels.h"

namespace sf {

// compute_add_sources_viscoelastic.F90:145-152: a(:,iglob) += sourcearray(isrc,:,i,j,k) * stf(isrc)
__global__ void k_add_sources_el(int n, const int* __restrict__ ig, const int* __restrict__ owner,
                                 const int* __restrict__ ploc, const realw* __restrict__ arr,
                                 const realw* __restrict__ stf, realw* accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  const int s = owner[t];
  const realw* sa = arr + ((int64_t)s * 125 + ploc[t]) * 3;
  const realw st = stf[s];
  accel[3 * g] = accel[3 * g] + sa[0] * st;
  accel[3 * g + 1] = accel[3 * g + 1] + sa[1] * st;
  accel[3 * g + 2] = accel[3 * g + 2] + sa[2] * st;
}

// compute_add_sources_acoustic.f90: chi'' -= sourcearray * stf / kappa (pressure source)
__global__ void k_add_sources_ac(int n, const int* __restrict__ ig, const int* __restrict__ owner,
                                 const int* __restrict__ ploc, const realw* __restrict__ arr,
                                 const realw* __restrict__ kap, const realw* __restrict__ stf, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int g = ig[t];
  const int s = owner[t];
  const realw sa = arr[((int64_t)s * 125 + ploc[t]) * 3];
  potdd[g] = potdd[g] - sa * stf[s] / kap[(int64_t)s * 125 + ploc[t]];
}

// compute_add_sources_viscoelastic.F90:226-245: a += adj(:,irec) * hxi heta hgamma (CUSTOM_REAL)
__global__ void k_adjoint_sources_el(int n, const int* __restrict__ ig, const int* __restrict__ owner,
                                     const int* __restrict__ ploc, const realw* __restrict__ h,
                                     const realw* __restrict__ adj, realw* accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  const int r = owner[t];
  const realw hl = h[(int64_t)r * 125 + ploc[t]];
  accel[3 * g] = accel[3 * g] + adj[3 * r] * hl;
  accel[3 * g + 1] = accel[3 * g + 1] + adj[3 * r + 1] * hl;
  accel[3 * g + 2] = accel[3 * g + 2] + adj[3 * r + 2] * hl;
}

// compute_add_sources_acoustic.f90:205-235 (adjoint sources in fluid elements):
// chi'' += adj(1, irec) * hxi heta hgamma, the product of the CUSTOM_REAL interpolants, added as is
// (no 1/kappa, positive sign)
__global__ void k_adjoint_sources_ac(int n, const int* __restrict__ ig, const int* __restrict__ owner,
                                     const int* __restrict__ ploc, const realw* __restrict__ h,
                                     const realw* __restrict__ adj, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int g = ig[t];
  const int r = owner[t];
  potdd[g] = potdd[g] + adj[3 * r] * h[(int64_t)r * 125 + ploc[t]];
}

// receiver interpolation (compute_interpolated_dva.f90:84-100): double-precision sum of
// h_i h_j h_k u over the 125 element points, rotated by nu_rec; one wave64 per listed receiver.
__global__ void __launch_bounds__(64) k_seismo_el(int n, const int* __restrict__ list, const int* __restrict__ ig,
                                                  const double* __restrict__ h, const double* __restrict__ nu,
                                                  const realw* __restrict__ field, realw* __restrict__ out) {
  if ((int)blockIdx.x >= n) return;
  const int r = list ? list[blockIdx.x] : (int)blockIdx.x;
  const int lane = threadIdx.x;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int p = lane; p < 125; p += 64) {
    const int64_t g = ig[(int64_t)r * 125 + p];
    const double hl = h[(int64_t)r * 125 + p];
    sx += (double)field[3 * g] * hl;
    sy += (double)field[3 * g + 1] * hl;
    sz += (double)field[3 * g + 2] * hl;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_xor(sx, off, 64);
    sy += __shfl_xor(sy, off, 64);
    sz += __shfl_xor(sz, off, 64);
  }
  if (lane == 0) {
    const double* n = nu + (int64_t)r * 9;
    out[3 * r] = (realw)(n[0] * sx + n[1] * sy + n[2] * sz);
    out[3 * r + 1] = (realw)(n[3] * sx + n[4] * sy + n[5] * sz);
    out[3 * r + 2] = (realw)(n[6] * sx + n[7] * sy + n[8] * sz);
  }
}

void launch_add_sources_el(hipStream_t s, const FacePoints& f, const realw* arr, const realw* stf, realw* accel) {
  for (int r = 0; r < f.nrounds; r++) {
    const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_add_sources_el, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o, f.ispec + o,
                       f.ploc + o, arr, stf, accel);
  }
  SF_LAUNCH_CHECK("k_add_sources_el");
}

void launch_add_sources_ac(hipStream_t s, const FacePoints& f, const realw* arr, const realw* kappa,
                           const realw* stf, realw* potdd) {
  for (int r = 0; r < f.nrounds; r++) {
    const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_add_sources_ac, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o, f.ispec + o,
                       f.ploc + o, arr, kappa, stf, potdd);
  }
  SF_LAUNCH_CHECK("k_add_sources_ac");
}

void launch_adjoint_sources_el(hipStream_t s, const FacePoints& f, const realw* h, const realw* adj, realw* accel) {
  for (int r = 0; r < f.nrounds; r++) {
    const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_adjoint_sources_el, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o, f.ispec + o,
                       f.ploc + o, h, adj, accel);
  }
  SF_LAUNCH_CHECK("k_adjoint_sources_el");
}

void launch_adjoint_sources_ac(hipStream_t s, const FacePoints& f, const realw* h, const realw* adj, realw* potdd) {
  for (int r = 0; r < f.nrounds; r++) {
    const int o = f.round_off[r], cnt = f.round_off[r + 1] - o;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_adjoint_sources_ac, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o, f.ispec + o,
                       f.ploc + o, h, adj, potdd);
  }
  SF_LAUNCH_CHECK("k_adjoint_sources_ac");
}

void launch_seismo_el(hipStream_t s, int n, const int* list, const int* iglob, const double* h, const double* nu,
                      const realw* field, realw* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_seismo_el, dim3(n), dim3(64), 0, s, n, list, iglob, h, nu, field, out);
  SF_LAUNCH_CHECK("k_seismo_el");
}

}  // namespace sf
