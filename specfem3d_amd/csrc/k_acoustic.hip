// k_acoustic.hip -- acoustic (fluid potential) element stiffness for NGLL=5.
//
// Restates compute_forces_acoustic.F90:35-726: chi'' -= div((1/rho) grad chi) in weak form.
// Same tiling as the elastic kernel (5 elements per 128-lane workgroup, one GLL column per
// lane, gamma contractions in registers, xi/eta through LDS). Compiled twice like
// k_elastic.hip (SF_EXACT = 1: no fused multiply-add, the reference's rounding).
#include "kernels.h"

#ifndef SF_EXACT
#define SF_EXACT 1
#endif
#if SF_EXACT
#define SFX(name) name##_exact
#define SF_VNS v_exact
#else
#define SFX(name) name##_fast
#define SF_VNS v_fast
#endif

namespace sf {
namespace SF_VNS {

template <int MODE>
__global__ void __launch_bounds__(128) k_acoustic(AcousticArgs A, GLLk G) {
  __shared__ realw s_buf[5][2][128];
  const int tid = threadIdx.x;
  const int el = tid / 25;
  const int c = tid - el * 25;
  const int i = c % 5, j = c / 5;
  const int ie = blockIdx.x * 5 + el;
  const bool active = (tid < 125) && (ie < A.nelem);
  const int e = active ? A.elems[ie] : 0;
  const int elx = el < 5 ? el : 4;

  int g[5];
  realw chi[5];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[k] = A.ibool[(int64_t)e * 125 + c + 25 * k];
      chi[k] = A.pot[g[k]];
      s_buf[el][0][c + 25 * k] = chi[k];
    }
  }
  __syncthreads();
  const int irr = active ? A.irregular[e] : 0;
  realw t1[5], t2[5], t3[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    if (!active) continue;
    const int p = c + 25 * k;
    const realw* s = s_buf[elx][0];
    const int b1 = 5 * j + 25 * k, b2 = i + 25 * k;
    realw a1 = G.hp[i * 5] * s[b1];
    realw a2 = G.hp[j * 5] * s[b2];
    realw a3 = G.hp[k * 5] * chi[0];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      a1 = a1 + G.hp[i * 5 + l] * s[b1 + l];
      a2 = a2 + G.hp[j * 5 + l] * s[b2 + 5 * l];
      a3 = a3 + G.hp[k * 5 + l] * chi[l];
    }
    const int64_t ep = (int64_t)e * 125 + p;
    const realw rho_invl = (realw)1 / A.rho[ep];
    if (irr) {
      const int64_t q = (int64_t)(irr - 1) * 125 + p;
      const int64_t P = A.jac_plane;
      const realw* J = A.jac;
      const realw xixl = J[q], xiyl = J[P + q], xizl = J[2 * P + q], etaxl = J[3 * P + q], etayl = J[4 * P + q];
      const realw etazl = J[5 * P + q], gaxl = J[6 * P + q], gayl = J[7 * P + q], gazl = J[8 * P + q];
      const realw jl = J[9 * P + q];
      const realw dpdx = xixl * a1 + etaxl * a2 + gaxl * a3;
      const realw dpdy = xiyl * a1 + etayl * a2 + gayl * a3;
      const realw dpdz = xizl * a1 + etazl * a2 + gazl * a3;
      t1[k] = rho_invl * jl * (xixl * dpdx + xiyl * dpdy + xizl * dpdz);
      t2[k] = rho_invl * jl * (etaxl * dpdx + etayl * dpdy + etazl * dpdz);
      t3[k] = rho_invl * jl * (gaxl * dpdx + gayl * dpdy + gazl * dpdz);
    } else {
      const realw jl = A.jacobian_regular, xr = A.xix_regular;
      t1[k] = rho_invl * jl * xr * xr * a1;
      t2[k] = rho_invl * jl * xr * xr * a2;
      t3[k] = rho_invl * jl * xr * xr * a3;
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      s_buf[el][0][c + 25 * k] = t1[k];
      s_buf[el][1][c + 25 * k] = t2[k];
    }
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int b1 = 5 * j + 25 * k, b2 = i + 25 * k;
    const realw* s1 = s_buf[el][0];
    const realw* s2 = s_buf[el][1];
    realw n1 = G.hpw[0 * 5 + i] * s1[b1];
    realw n2 = G.hpw[0 * 5 + j] * s2[b2];
    realw n3 = G.hpw[0 * 5 + k] * t3[0];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      n1 = n1 + G.hpw[l * 5 + i] * s1[b1 + l];
      n2 = n2 + G.hpw[l * 5 + j] * s2[b2 + 5 * l];
      n3 = n3 + G.hpw[l * 5 + k] * t3[l];
    }
    const realw f = G.wxy[j * 5 + k] * n1 + G.wxy[i * 5 + k] * n2 + G.wxy[i * 5 + j] * n3;
    if (MODE == 1) {
      A.local_out[(int64_t)e * 125 + c + 25 * k] = f;
    } else if (MODE == 2) {
      A.potdd[g[k]] = A.potdd[g[k]] + (-f);
    } else {
      atomicAdd(A.potdd + g[k], -f);
    }
  }
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_acoustic)(hipStream_t s, const AcousticArgs& a, int mode) {
  if (a.nelem <= 0) return;
  dim3 grid((a.nelem + 4) / 5), block(128);
  if (mode == 1)
    hipLaunchKernelGGL(k_acoustic<1>, grid, block, 0, s, a, g_gll_host);
  else if (mode == 2)
    hipLaunchKernelGGL(k_acoustic<2>, grid, block, 0, s, a, g_gll_host);
  else
    hipLaunchKernelGGL(k_acoustic<0>, grid, block, 0, s, a, g_gll_host);
  SF_LAUNCH_CHECK("k_acoustic");
}

}  // namespace sf
