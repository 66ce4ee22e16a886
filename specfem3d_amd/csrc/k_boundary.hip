// k_boundary.hip -- face-point kernels: Stacey absorbing boundaries, fluid-solid coupling,
// acoustic free surface.
//
// Faces are flattened at prepare time into face points carrying the global index, normal,
// weight and impedances, and sorted into "rounds" (the r-th occurrence of a global point in
// the host's face order). One lane per face point; one launch per round updates with plain
// read-modify-writes (no two lanes of a round touch the same point), so the sums at points
// shared by several faces are accumulated in the reference's face order and are
// deterministic. When the element stiffness itself scatters with float atomics (so the sums
// have no fixed order anyway), the rounds run as one launch with atomic adds (ATOMIC).
// Built with -ffp-contract=off (reference rounding).
#include "kernels.h"

namespace sf {

// Stacey elastic (compute_stacey_viscoelastic.F90:85-118):
//   t = rho vp (v.n) n + rho vs (v - (v.n) n);  a -= t w
// `store` (optional) receives t w at the point's record slot (SAVE_FORWARD boundary record).
// Float atomics (atomic assembly with the elastic Stacey term in one launch).
__global__ void k_stacey_el(int n, const int* __restrict__ ig, const realw* __restrict__ nrm,
                            const realw* __restrict__ jw, const realw* __restrict__ rvp,
                            const realw* __restrict__ rvs, const int* __restrict__ slot,
                            const realw* __restrict__ veloc, realw* accel, realw* __restrict__ store) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  const realw vx = veloc[3 * g], vy = veloc[3 * g + 1], vz = veloc[3 * g + 2];
  const realw nx = nrm[3 * t], ny = nrm[3 * t + 1], nz = nrm[3 * t + 2];
  const realw vn = vx * nx + vy * ny + vz * nz;
  const realw p = rvp[t], s = rvs[t];
  const realw tx = p * vn * nx + s * (vx - vn * nx);
  const realw ty = p * vn * ny + s * (vy - vn * ny);
  const realw tz = p * vn * nz + s * (vz - vn * nz);
  const realw w = jw[t];
  const realw fx = tx * w, fy = ty * w, fz = tz * w;
  atomicAdd(accel + 3 * g, -fx);
  atomicAdd(accel + 3 * g + 1, -fy);
  atomicAdd(accel + 3 * g + 2, -fz);
  if (store) {
    const int q = slot[t];
    store[3 * q] = fx;
    store[3 * q + 1] = fy;
    store[3 * q + 2] = fz;
  }
}

// The same term in one launch without atomics: a lane per point of round 0 (the first occurrence of
// each global point) walks the point's occurrence chain in round order and applies each face's term
// to a register copy of accel, which is the read-modify-write sequence of the rounds, in their
// order, so the sum has the same bits.
__global__ void k_stacey_el_chain(int n0, const int* __restrict__ ig, const realw* __restrict__ nrm,
                                  const realw* __restrict__ jw, const realw* __restrict__ rvp,
                                  const realw* __restrict__ rvs, const int* __restrict__ slot,
                                  const int* __restrict__ next, const realw* __restrict__ veloc, realw* accel,
                                  realw* __restrict__ store) {
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= n0) return;
  const int64_t g = ig[t0];
  const realw vx = veloc[3 * g], vy = veloc[3 * g + 1], vz = veloc[3 * g + 2];
  realw ax = accel[3 * g], ay = accel[3 * g + 1], az = accel[3 * g + 2];
  for (int t = t0; t >= 0; t = next[t]) {
    const realw nx = nrm[3 * t], ny = nrm[3 * t + 1], nz = nrm[3 * t + 2];
    const realw vn = vx * nx + vy * ny + vz * nz;
    const realw p = rvp[t], s = rvs[t];
    const realw tx = p * vn * nx + s * (vx - vn * nx);
    const realw ty = p * vn * ny + s * (vy - vn * ny);
    const realw tz = p * vn * nz + s * (vz - vn * nz);
    const realw w = jw[t];
    const realw fx = tx * w, fy = ty * w, fz = tz * w;
    ax = ax - fx;
    ay = ay - fy;
    az = az - fz;
    if (store) {
      const int q = slot[t];
      store[3 * q] = fx;
      store[3 * q + 1] = fy;
      store[3 * q + 2] = fz;
    }
  }
  accel[3 * g] = ax;
  accel[3 * g + 1] = ay;
  accel[3 * g + 2] = az;
}

// replay of a stored boundary record on the backward field (compute_stacey_viscoelastic.F90:274-283)
__global__ void k_stacey_el_replay(int n, const int* __restrict__ ig, const int* __restrict__ slot,
                                   const realw* __restrict__ rec, realw* b_accel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  const int q = slot[t];
  b_accel[3 * g] = b_accel[3 * g] - rec[3 * q];
  b_accel[3 * g + 1] = b_accel[3 * g + 1] - rec[3 * q + 1];
  b_accel[3 * g + 2] = b_accel[3 * g + 2] - rec[3 * q + 2];
}

// Stacey acoustic (compute_stacey_acoustic.f90:96-112): chi'' -= chi' w / c / rho.
// imp_p holds rho, imp_s holds c = sqrt(kappa/rho), both per face point.
template <bool ATOMIC>
__global__ void k_stacey_ac(int n, const int* __restrict__ ig, const realw* __restrict__ jw,
                            const realw* __restrict__ rho, const realw* __restrict__ cp, const int* __restrict__ slot,
                            const realw* __restrict__ potd, realw* potdd, realw* __restrict__ store) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int g = ig[t];
  const realw absorbl = potd[g] * jw[t] / cp[t] / rho[t];
  if (ATOMIC) atomicAdd(potdd + g, -absorbl);
  else potdd[g] = potdd[g] - absorbl;
  if (store) store[slot[t]] = absorbl;
}

__global__ void k_stacey_ac_replay(int n, const int* __restrict__ ig, const int* __restrict__ slot,
                                   const realw* __restrict__ rec, realw* b_potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int g = ig[t];
  b_potdd[g] = b_potdd[g] - rec[slot[t]];
}

// acoustic side of the coupling (compute_coupling_acoustic_el.f90:84-151): chi'' += w (u.n)
template <bool ATOMIC>
__global__ void k_coupling_ac_el(int n, const int* __restrict__ ig, const realw* __restrict__ nrm,
                                 const realw* __restrict__ jw, const realw* __restrict__ displ, realw* potdd,
                                 int flip) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  realw ux = displ[3 * g], uy = displ[3 * g + 1], uz = displ[3 * g + 2];
  if (flip) { ux = -ux; uy = -uy; uz = -uz; }
  const realw un = ux * nrm[3 * t] + uy * nrm[3 * t + 1] + uz * nrm[3 * t + 2];
  if (ATOMIC) atomicAdd(potdd + g, jw[t] * un);
  else potdd[g] = potdd[g] + jw[t] * un;
}

// elastic side (compute_coupling_viscoelastic_ac.f90:84-160): a += w n (-chi'')
template <bool ATOMIC>
__global__ void k_coupling_el_ac(int n, const int* __restrict__ ig, const realw* __restrict__ nrm,
                                 const realw* __restrict__ jw, const realw* __restrict__ potdd, realw* accel,
                                 int flip) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = ig[t];
  realw pr = -potdd[g];
  if (flip) pr = -pr;
  const realw w = jw[t];
  if (ATOMIC) {
    atomicAdd(accel + 3 * g, w * nrm[3 * t] * pr);
    atomicAdd(accel + 3 * g + 1, w * nrm[3 * t + 1] * pr);
    atomicAdd(accel + 3 * g + 2, w * nrm[3 * t + 2] * pr);
  } else {
    accel[3 * g] = accel[3 * g] + w * nrm[3 * t] * pr;
    accel[3 * g + 1] = accel[3 * g + 1] + w * nrm[3 * t + 1] * pr;
    accel[3 * g + 2] = accel[3 * g + 2] + w * nrm[3 * t + 2] * pr;
  }
}

__global__ void k_free_surface(int n, const int* __restrict__ ig, realw* pot, realw* potd, realw* potdd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int g = ig[t];
  pot[g] = 0.f;
  potd[g] = 0.f;
  potdd[g] = 0.f;
}

#define SF_ROUNDS(f, body)                                   \
  for (int r = 0; r < (f).nrounds; r++) {                    \
    const int o = (f).round_off[r];                          \
    const int cnt = (f).round_off[r + 1] - o;                \
    if (cnt <= 0) continue;                                  \
    body;                                                    \
  }

void launch_stacey_elastic(hipStream_t s, const FacePoints& f, const realw* veloc, realw* accel, realw* b_absorb,
                           int mode, bool atomic) {
  if (atomic) {
    if (f.n > 0)
      hipLaunchKernelGGL(k_stacey_el, sf_grid(f.n, 128), dim3(128), 0, s, f.n, f.iglob, f.normal, f.jacw,
                         f.imp_p, f.imp_s, f.slot, veloc, accel, mode ? b_absorb : (realw*)nullptr);
  } else if (f.nrounds > 0 && f.round_off[1] > 0) {
    const int n0 = f.round_off[1];
    hipLaunchKernelGGL(k_stacey_el_chain, sf_grid(n0, 128), dim3(128), 0, s, n0, f.iglob, f.normal, f.jacw, f.imp_p,
                       f.imp_s, f.slot, f.next, veloc, accel, mode ? b_absorb : (realw*)nullptr);
  }
  SF_LAUNCH_CHECK("k_stacey_el");
}

void launch_stacey_elastic_replay(hipStream_t s, const FacePoints& f, const realw* rec, realw* b_accel) {
  SF_ROUNDS(f, hipLaunchKernelGGL(k_stacey_el_replay, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o,
                                  f.slot + o, rec, b_accel));
  SF_LAUNCH_CHECK("k_stacey_el_replay");
}

void launch_stacey_acoustic(hipStream_t s, const FacePoints& f, const realw* potd, realw* potdd, realw* b_absorb,
                            int mode, bool atomic) {
  if (atomic) {
    if (f.n > 0)
      hipLaunchKernelGGL(k_stacey_ac<true>, sf_grid(f.n, 128), dim3(128), 0, s, f.n, f.iglob, f.jacw, f.imp_p,
                         f.imp_s, f.slot, potd, potdd, mode ? b_absorb : (realw*)nullptr);
  } else {
    SF_ROUNDS(f, hipLaunchKernelGGL(k_stacey_ac<false>, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o,
                                    f.jacw + o, f.imp_p + o, f.imp_s + o, f.slot + o, potd, potdd,
                                    mode ? b_absorb : (realw*)nullptr));
  }
  SF_LAUNCH_CHECK("k_stacey_ac");
}

void launch_stacey_acoustic_replay(hipStream_t s, const FacePoints& f, const realw* rec, realw* b_potdd) {
  SF_ROUNDS(f, hipLaunchKernelGGL(k_stacey_ac_replay, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o,
                                  f.slot + o, rec, b_potdd));
  SF_LAUNCH_CHECK("k_stacey_ac_replay");
}

void launch_coupling_ac_el(hipStream_t s, const FacePoints& f, const realw* displ, realw* potdd, int flip,
                           bool atomic) {
  if (atomic) {
    if (f.n > 0)
      hipLaunchKernelGGL(k_coupling_ac_el<true>, sf_grid(f.n, 128), dim3(128), 0, s, f.n, f.iglob, f.normal, f.jacw,
                         displ, potdd, flip);
  } else {
    SF_ROUNDS(f, hipLaunchKernelGGL(k_coupling_ac_el<false>, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o,
                                    f.normal + 3 * o, f.jacw + o, displ, potdd, flip));
  }
  SF_LAUNCH_CHECK("k_coupling_ac_el");
}

void launch_coupling_el_ac(hipStream_t s, const FacePoints& f, const realw* potdd, realw* accel, int flip,
                           bool atomic) {
  if (atomic) {
    if (f.n > 0)
      hipLaunchKernelGGL(k_coupling_el_ac<true>, sf_grid(f.n, 128), dim3(128), 0, s, f.n, f.iglob, f.normal, f.jacw,
                         potdd, accel, flip);
  } else {
    SF_ROUNDS(f, hipLaunchKernelGGL(k_coupling_el_ac<false>, sf_grid(cnt, 128), dim3(128), 0, s, cnt, f.iglob + o,
                                    f.normal + 3 * o, f.jacw + o, potdd, accel, flip));
  }
  SF_LAUNCH_CHECK("k_coupling_el_ac");
}

void launch_free_surface(hipStream_t s, const FacePoints& f, realw* pot, realw* potd, realw* potdd) {
  if (f.n <= 0) return;
  hipLaunchKernelGGL(k_free_surface, sf_grid(f.n, 128), dim3(128), 0, s, f.n, f.iglob, pot, potd, potdd);
  SF_LAUNCH_CHECK("k_free_surface");
}

}  // namespace sf
