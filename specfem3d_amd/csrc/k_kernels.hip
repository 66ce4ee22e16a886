// k_kernels.hip -- sensitivity-kernel accumulation, ordered assembly gather, max-norm.
// Built with -ffp-contract=off (reference rounding).
#include "kernels.h"

namespace sf {

// compute_kernels.f90:151-186 (isotropic), one lane per local point of the elastic elements:
//   rho_kl   += dt (a . b_u)
//   mu_kl    += dt (e1 b1 + e2 b2 + (e1+e2)(b1+b2) + 2 (e3 b3 + e4 b4 + e5 b5))
//   kappa_kl += dt (9 tr/3 b_tr/3)
__global__ void k_compute_kernels_el(int nelem, const int* __restrict__ elems, const int* __restrict__ ibool,
                                     const realw* __restrict__ accel, const realw* __restrict__ b_displ,
                                     const realw* __restrict__ eps, const realw* __restrict__ etr3,
                                     const realw* __restrict__ beps, const realw* __restrict__ btr3, int64_t nl,
                                     realw deltat, realw* rho_kl, realw* mu_kl, realw* kappa_kl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nelem * 125) return;
  const int e = elems[t / 125];
  const int p = (int)(t % 125);
  const int64_t ep = (int64_t)e * 125 + p;
  const int64_t g = ibool[ep];
  const realw dot = accel[3 * g] * b_displ[3 * g] + accel[3 * g + 1] * b_displ[3 * g + 1] +
                    accel[3 * g + 2] * b_displ[3 * g + 2];
  rho_kl[ep] = rho_kl[ep] + deltat * dot;
  const realw e1 = eps[eps_idx(e, 0, p)], e2 = eps[eps_idx(e, 1, p)], e3 = eps[eps_idx(e, 2, p)],
              e4 = eps[eps_idx(e, 3, p)], e5 = eps[eps_idx(e, 4, p)];
  const realw b1 = beps[eps_idx(e, 0, p)], b2 = beps[eps_idx(e, 1, p)], b3 = beps[eps_idx(e, 2, p)],
              b4 = beps[eps_idx(e, 3, p)], b5 = beps[eps_idx(e, 4, p)];
  mu_kl[ep] = mu_kl[ep] + deltat * (e1 * b1 + e2 * b2 + (e1 + e2) * (b1 + b2) + (realw)2 * (e3 * b3 + e4 * b4 + e5 * b5));
  kappa_kl[ep] = kappa_kl[ep] + deltat * ((realw)9 * etr3[ep] * btr3[ep]);
}

// compute_kernels.f90:257-337 (compute_kernels_ac), one workgroup per fluid element:
//   rho_ac_kl   += dt (grad chi_adj . grad chi_b)   (gradients as compute_gradient_in_acoustic:
//                 contractions, inverse Jacobian, 1/rho, then times rho again)
//   kappa_ac_kl += dt chi_adj chi_b''
__device__ __forceinline__ void ac_gradient(const realw* s, const realw* hp, int i, int j, int k, int irr,
                                            const realw* jac, int64_t P, int64_t q, realw xr, realw rho_invl,
                                            realw* g) {
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
  realw t1 = hp[i * 5] * s[b1], t2 = hp[j * 5] * s[b2], t3 = hp[k * 5] * s[b3];
#pragma unroll
  for (int l = 1; l < 5; l++) {
    t1 = t1 + hp[i * 5 + l] * s[b1 + l];
    t2 = t2 + hp[j * 5 + l] * s[b2 + 5 * l];
    t3 = t3 + hp[k * 5 + l] * s[b3 + 25 * l];
  }
  if (irr) {
    const realw xixl = jac[q], xiyl = jac[P + q], xizl = jac[2 * P + q], etaxl = jac[3 * P + q];
    const realw etayl = jac[4 * P + q], etazl = jac[5 * P + q], gaxl = jac[6 * P + q], gayl = jac[7 * P + q];
    const realw gazl = jac[8 * P + q];
    g[0] = (t1 * xixl + t2 * etaxl + t3 * gaxl) * rho_invl;
    g[1] = (t1 * xiyl + t2 * etayl + t3 * gayl) * rho_invl;
    g[2] = (t1 * xizl + t2 * etazl + t3 * gazl) * rho_invl;
  } else {
    g[0] = t1 * xr * rho_invl;
    g[1] = t2 * xr * rho_invl;
    g[2] = t3 * xr * rho_invl;
  }
}

__global__ void __launch_bounds__(128) k_compute_kernels_ac(int nelem, const int* __restrict__ elems,
                                                            const int* __restrict__ ibool,
                                                            const int* __restrict__ irregular, realw xix_regular,
                                                            const realw* __restrict__ jac, int64_t jac_plane,
                                                            const realw* __restrict__ rhostore,
                                                            const realw* __restrict__ pot,
                                                            const realw* __restrict__ b_pot,
                                                            const realw* __restrict__ b_potdd, GLLk G, realw deltat,
                                                            realw* rho_ac_kl, realw* kappa_ac_kl) {
  __shared__ realw sb[125], sa[125], hp[25];
  const int p = threadIdx.x;
  if (p < 25) hp[p] = G.hp[p];
  const int e = elems[blockIdx.x];
  const int64_t ep = (int64_t)e * 125 + (p < 125 ? p : 124);
  const int g = ibool[ep];
  if (p < 125) {
    sb[p] = b_pot[g];
    sa[p] = pot[g];
  }
  __syncthreads();
  if (p >= 125) return;
  const int k = p / 25, j = (p / 5) % 5, i = p % 5;
  const int irr = irregular[e];
  const int64_t q = (int64_t)(irr > 0 ? irr - 1 : 0) * 125 + p;
  const realw rhol = rhostore[ep];
  const realw rho_invl = (realw)1 / rhol;
  realw bg[3], ag[3];
  ac_gradient(sb, hp, i, j, k, irr, jac, jac_plane, q, xix_regular, rho_invl, bg);
  ac_gradient(sa, hp, i, j, k, irr, jac, jac_plane, q, xix_regular, rho_invl, ag);
  const realw a1 = rhol * ag[0], a2 = rhol * ag[1], a3 = rhol * ag[2];
  const realw b1 = rhol * bg[0], b2 = rhol * bg[1], b3 = rhol * bg[2];
  rho_ac_kl[ep] = rho_ac_kl[ep] + deltat * (a1 * b1 + a2 * b2 + a3 * b3);
  kappa_ac_kl[ep] = kappa_ac_kl[ep] + deltat * sa[p] * b_potdd[g];
}

// receivers in acoustic elements (compute_seismograms.f90:127-144, compute_interpolated_dva_acoust
// compute_interpolated_dva.f90:219-328), one workgroup per listed receiver, FP64 sums like the CPU path:
//   pressure: p = -sum_ijk h_ijk field(ibool)   (field = potential with the pressure trick, else
//             potential_dot_dot); stored in component 0, components 1 and 2 zero;
//   vector:   the receiver element's (1/rho) grad(field) at its GLL points
//             (compute_gradient_in_acoustic_fast_Deville), interpolated and rotated by nu_rec.
__global__ void __launch_bounds__(128) k_seismo_ac(AcReceiverArgs A, GLLk G, const realw* __restrict__ field,
                                                   int pressure, realw* __restrict__ out) {
  __shared__ realw s[125], hp[25];
  __shared__ double red[2][3];
  const int p = threadIdx.x;
  const int r = A.list ? A.list[blockIdx.x] : (int)blockIdx.x;
  if (p < 25) hp[p] = G.hp[p];
  const int q = p < 125 ? p : 124;
  const int g = A.iglob[(int64_t)r * 125 + q];
  if (p < 125) s[p] = field[g];
  __syncthreads();
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (p < 125) {
    const double hl = A.h[(int64_t)r * 125 + p];
    if (pressure) {
      v0 = -(hl * (double)s[p]);
    } else {
      const int e = A.rec_ispec[r];
      const int k = p / 25, j = (p / 5) % 5, i = p % 5;
      const int irr = A.irregular[e];
      const int64_t qq = (int64_t)(irr > 0 ? irr - 1 : 0) * 125 + p;
      const realw rho_invl = (realw)1 / A.rho[(int64_t)e * 125 + p];
      realw gv[3];
      ac_gradient(s, hp, i, j, k, irr, A.jac, A.jac_plane, qq, A.xix_regular, rho_invl, gv);
      v0 = hl * (double)gv[0];
      v1 = hl * (double)gv[1];
      v2 = hl * (double)gv[2];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
    v2 += __shfl_xor(v2, off, 64);
  }
  if ((p & 63) == 0) {
    red[p >> 6][0] = v0;
    red[p >> 6][1] = v1;
    red[p >> 6][2] = v2;
  }
  __syncthreads();
  if (p != 0) return;
  const double x = red[0][0] + red[1][0], y = red[0][1] + red[1][1], z = red[0][2] + red[1][2];
  if (pressure) {
    out[3 * r] = (realw)x;
    out[3 * r + 1] = 0.f;
    out[3 * r + 2] = 0.f;
  } else {
    const double* n = A.nu + (int64_t)r * 9;
    out[3 * r] = (realw)(n[0] * x + n[1] * y + n[2] * z);
    out[3 * r + 1] = (realw)(n[3] * x + n[4] * y + n[5] * z);
    out[3 * r + 2] = (realw)(n[6] * x + n[7] * y + n[8] * z);
  }
}

// ordered assembly: for each global point, subtract its element contributions in increasing
// element order (the order of the reference's serial element loop). ptr/idx is a CSR list of
// local slots e*125+p per global point, sorted by element.
__global__ void k_ordered_gather(int nglob, int ncomp, const int* __restrict__ ptr, const int* __restrict__ idx,
                                 const realw* __restrict__ local, int64_t nl, realw* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  const int b = ptr[g], e = ptr[g + 1];
  if (b == e) return;
  for (int c = 0; c < ncomp; c++) {
    realw a = out[(int64_t)g * ncomp + c];
    for (int q = b; q < e; q++) a = a - local[c * nl + idx[q]];
    out[(int64_t)g * ncomp + c] = a;
  }
}

// max over points of the vector norm (check_stability.f90); NaN propagates as +inf bits
__global__ void k_max_norm(int64_t npts, int ncomp, const realw* __restrict__ f, unsigned int* out) {
  __shared__ realw sm[256];
  realw m = 0.f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (int64_t)gridDim.x * blockDim.x) {
    realw v;
    if (ncomp == 3) {
      const realw x = f[3 * t], y = f[3 * t + 1], z = f[3 * t + 2];
      v = sqrtf(x * x + y * y + z * z);
    } else {
      v = fabsf(f[t]);
    }
    if (v != v) v = INFINITY;
    m = fmaxf(m, v);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(sm[0]));  // non-negative floats order as uints
}

__global__ void k_zero_u32(unsigned int* p) { *p = 0u; }

void launch_compute_kernels_el(hipStream_t s, int nelem, const int* elems, const int* ibool, const realw* accel,
                               const realw* b_displ, const realw* eps, const realw* eps_tr3, const realw* b_eps,
                               const realw* b_eps_tr3, int64_t nloc, realw deltat, realw* rho_kl, realw* mu_kl,
                               realw* kappa_kl) {
  if (nelem <= 0) return;
  hipLaunchKernelGGL(k_compute_kernels_el, sf_grid((int64_t)nelem * 125, 256), dim3(256), 0, s, nelem, elems, ibool,
                     accel, b_displ, eps, eps_tr3, b_eps, b_eps_tr3, nloc, deltat, rho_kl, mu_kl, kappa_kl);
  SF_LAUNCH_CHECK("k_compute_kernels_el");
}

void launch_compute_kernels_ac(hipStream_t s, int nelem, const int* elems, const int* ibool, const int* irregular,
                               realw xix_regular, const realw* jac, int64_t jac_plane, const realw* rhostore,
                               const realw* pot, const realw* b_pot, const realw* b_potdd, realw deltat,
                               realw* rho_ac_kl, realw* kappa_ac_kl) {
  if (nelem <= 0) return;
  hipLaunchKernelGGL(k_compute_kernels_ac, dim3((unsigned)nelem), dim3(128), 0, s, nelem, elems, ibool, irregular,
                     xix_regular, jac, jac_plane, rhostore, pot, b_pot, b_potdd, g_gll_host, deltat, rho_ac_kl,
                     kappa_ac_kl);
  SF_LAUNCH_CHECK("k_compute_kernels_ac");
}

void launch_seismo_ac(hipStream_t s, const AcReceiverArgs& a, const realw* field, int pressure, realw* out) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_seismo_ac, dim3((unsigned)a.n), dim3(128), 0, s, a, g_gll_host, field, pressure, out);
  SF_LAUNCH_CHECK("k_seismo_ac");
}

void launch_ordered_gather(hipStream_t s, int nglob, int ncomp, const int* ptr, const int* idx, const realw* local,
                           int64_t nloc, realw* out) {
  hipLaunchKernelGGL(k_ordered_gather, sf_grid(nglob, 256), dim3(256), 0, s, nglob, ncomp, ptr, idx, local, nloc,
                     out);
  SF_LAUNCH_CHECK("k_ordered_gather");
}

void launch_max_norm(hipStream_t s, int64_t npts, int ncomp, const realw* f, realw* out) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, reinterpret_cast<unsigned int*>(out));
  int64_t nb = (npts + 255) / 256;
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_max_norm, dim3((unsigned)nb), dim3(256), 0, s, npts, ncomp, f,
                     reinterpret_cast<unsigned int*>(out));
  SF_LAUNCH_CHECK("k_max_norm");
}

}  // namespace sf
