// k_update.hip -- Newmark time-marching streams (pure HBM-bound, float4 vectorised).
//
// predictor   (update_displacement_scheme.f90:359-410, :206-227):
//     u += dt v + dt^2/2 a ; v += dt/2 a ; a = 0
// corrector   (compute_forces_viscoelastic_calling_routine.F90:368-374 + update_veloc_elastic):
//     a *= rmass^-1 ; v += dt/2 a
// corrector+predictor fused: the corrector of step n followed by the predictor of step n+1
// in one pass (same arithmetic, one read and one write of u, v, a instead of two).
// All arrays are flat (3*nglob for vector fields, nglob for potentials) and the inverse
// mass is stored with the same layout, so each thread handles 4 consecutive floats.
// Built with -ffp-contract=off: identical rounding to the CPU reference.
#include "kernels.h"

namespace sf {

__global__ void k_predictor(int64_t n4, int64_t n, realw* __restrict__ u, realw* __restrict__ v,
                            realw* __restrict__ a, realw dt, realw dt2, realw dtsq2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n4) {
    float4 U = reinterpret_cast<float4*>(u)[t];
    float4 V = reinterpret_cast<float4*>(v)[t];
    const float4 A = reinterpret_cast<float4*>(a)[t];
    U.x = U.x + dt * V.x + dtsq2 * A.x; V.x = V.x + dt2 * A.x;
    U.y = U.y + dt * V.y + dtsq2 * A.y; V.y = V.y + dt2 * A.y;
    U.z = U.z + dt * V.z + dtsq2 * A.z; V.z = V.z + dt2 * A.z;
    U.w = U.w + dt * V.w + dtsq2 * A.w; V.w = V.w + dt2 * A.w;
    reinterpret_cast<float4*>(u)[t] = U;
    reinterpret_cast<float4*>(v)[t] = V;
    reinterpret_cast<float4*>(a)[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // scalar tail
  const int64_t r = 4 * n4 + t;
  if (t < 4 && r < n) {
    u[r] = u[r] + dt * v[r] + dtsq2 * a[r];
    v[r] = v[r] + dt2 * a[r];
    a[r] = 0.f;
  }
}

__global__ void k_corrector(int64_t n4, int64_t n, realw* __restrict__ v, realw* __restrict__ a,
                            const realw* __restrict__ rm, realw dt2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n4) {
    float4 V = reinterpret_cast<float4*>(v)[t];
    float4 A = reinterpret_cast<float4*>(a)[t];
    const float4 M = reinterpret_cast<const float4*>(rm)[t];
    A.x = A.x * M.x; A.y = A.y * M.y; A.z = A.z * M.z; A.w = A.w * M.w;
    V.x = V.x + dt2 * A.x; V.y = V.y + dt2 * A.y; V.z = V.z + dt2 * A.z; V.w = V.w + dt2 * A.w;
    reinterpret_cast<float4*>(v)[t] = V;
    reinterpret_cast<float4*>(a)[t] = A;
  }
  const int64_t r = 4 * n4 + t;
  if (t < 4 && r < n) {
    a[r] = a[r] * rm[r];
    v[r] = v[r] + dt2 * a[r];
  }
}

// OLD = 1 also stores PML_displ_old = u + t1 v + t2 a of the corrected field before the predictor
// (the arithmetic of k_pml_store_old), for every point of the range
template <int OLD>
__global__ void k_corrector_predictor(int64_t n4, int64_t n, realw* __restrict__ u, realw* __restrict__ v,
                                      realw* __restrict__ a, const realw* __restrict__ rm, realw dt2p, realw dt,
                                      realw dt2, realw dtsq2, realw* __restrict__ old, realw t1, realw t2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n4) {
    float4 U = reinterpret_cast<float4*>(u)[t];
    float4 V = reinterpret_cast<float4*>(v)[t];
    float4 A = reinterpret_cast<float4*>(a)[t];
    const float4 M = reinterpret_cast<const float4*>(rm)[t];
    float4 O;
#define SF_CP(c)                               \
  A.c = A.c * M.c;                             \
  V.c = V.c + dt2p * A.c;                      \
  if (OLD) O.c = U.c + t1 * V.c + t2 * A.c;    \
  U.c = U.c + dt * V.c + dtsq2 * A.c;          \
  V.c = V.c + dt2 * A.c;
    SF_CP(x) SF_CP(y) SF_CP(z) SF_CP(w)
#undef SF_CP
    reinterpret_cast<float4*>(u)[t] = U;
    reinterpret_cast<float4*>(v)[t] = V;
    reinterpret_cast<float4*>(a)[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (OLD) reinterpret_cast<float4*>(old)[t] = O;
  }
  const int64_t r = 4 * n4 + t;
  if (t < 4 && r < n) {
    realw ar = a[r] * rm[r];
    realw vr = v[r] + dt2p * ar;
    if (OLD) old[r] = u[r] + t1 * vr + t2 * ar;
    u[r] = u[r] + dt * vr + dtsq2 * ar;
    v[r] = vr + dt2 * ar;
    a[r] = 0.f;
  }
}

// PML_displ_old (update_displ_elastic_PML, update_displacement_scheme.f90:417-455) at the global points
// of the C-PML elements, from the field before the predictor: u + t1 v + t2 a with
// t1 = deltatover2 (1 - 2 theta), t2 = deltatsqover2 (1 - theta). With a corrector still pending (it
// is fused into the predictor that follows), the corrected v and a are formed first with the
// corrector's own expressions.
__global__ void k_pml_store_old(int n, const int* __restrict__ pts, const realw* __restrict__ u,
                                const realw* __restrict__ v, const realw* __restrict__ a,
                                const realw* __restrict__ rm, int pending, realw dt2p, realw t1, realw t2,
                                realw* __restrict__ old) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = pts[t];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    realw ac = a[3 * g + c], vc = v[3 * g + c];
    if (pending) {
      ac = ac * rm[3 * g + c];
      vc = vc + dt2p * ac;
    }
    old[3 * g + c] = u[3 * g + c] + t1 * vc + t2 * ac;
  }
}

// Dirichlet condition on the outer faces of the C-PML layers (pml_impose_boundary_condition_elastic)
__global__ void k_zero_points(int n, const int* __restrict__ pts, realw* __restrict__ u, realw* __restrict__ v,
                              realw* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t g = pts[t];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    u[3 * g + c] = 0.f;
    v[3 * g + c] = 0.f;
    a[3 * g + c] = 0.f;
  }
}

static dim3 grid_for(int64_t n4, int block) {
  int64_t b = (n4 + block - 1) / block;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

void launch_predictor(hipStream_t s, int64_t n, realw* u, realw* v, realw* a, realw dt, realw dt2, realw dtsq2) {
  if (n <= 0) return;
  const int64_t n4 = n / 4;
  hipLaunchKernelGGL(k_predictor, grid_for(n4, 256), dim3(256), 0, s, n4, n, u, v, a, dt, dt2, dtsq2);
  SF_LAUNCH_CHECK("k_predictor");
}

void launch_corrector(hipStream_t s, int64_t n, realw* v, realw* a, const realw* rmass, realw dt2) {
  if (n <= 0) return;
  const int64_t n4 = n / 4;
  hipLaunchKernelGGL(k_corrector, grid_for(n4, 256), dim3(256), 0, s, n4, n, v, a, rmass, dt2);
  SF_LAUNCH_CHECK("k_corrector");
}

void launch_corrector_predictor(hipStream_t s, int64_t n, realw* u, realw* v, realw* a, const realw* rmass,
                                realw dt2_prev, realw dt, realw dt2, realw dtsq2, realw* old, realw t1, realw t2) {
  if (n <= 0) return;
  const int64_t n4 = n / 4;
  if (old)
    hipLaunchKernelGGL(k_corrector_predictor<1>, grid_for(n4, 256), dim3(256), 0, s, n4, n, u, v, a, rmass,
                       dt2_prev, dt, dt2, dtsq2, old, t1, t2);
  else
    hipLaunchKernelGGL(k_corrector_predictor<0>, grid_for(n4, 256), dim3(256), 0, s, n4, n, u, v, a, rmass,
                       dt2_prev, dt, dt2, dtsq2, nullptr, 0.f, 0.f);
  SF_LAUNCH_CHECK("k_corrector_predictor");
}

void launch_pml_store_old(hipStream_t s, int n, const int* pts, const realw* u, const realw* v, const realw* a,
                          const realw* rmass, int pending, realw dt2p, realw t1, realw t2, realw* old) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pml_store_old, grid_for(n, 256), dim3(256), 0, s, n, pts, u, v, a, rmass, pending, dt2p, t1,
                     t2, old);
  SF_LAUNCH_CHECK("k_pml_store_old");
}

void launch_zero_points(hipStream_t s, int n, const int* pts, realw* u, realw* v, realw* a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_points, grid_for(n, 256), dim3(256), 0, s, n, pts, u, v, a);
  SF_LAUNCH_CHECK("k_zero_points");
}

}  // namespace sf
