// k_acoustic.hip -- acoustic (fluid potential) element stiffness for NGLL=5.
//
// Restates compute_forces_acoustic.F90:35-726: chi'' -= div((1/rho) grad chi) in weak form.
// Point-per-lane tiling like k_elastic_pt (one element per 128-lane workgroup, the element's
// tiles in LDS). Compiled twice like k_elastic.hip (SF_EXACT = 1: no fused multiply-add, the
// reference's rounding).
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

// One element per 128-lane workgroup; lane p owns GLL point p = i + 5 j + 25 k (lanes 125..127
// shadow point 124 and store nothing). The potential tile and the three products go through LDS;
// every sum runs in the reference's order (the five terms of each contraction left to right, then
// the three weighted divergence terms), the exact build's rounding of the CPU solver.
__device__ __forceinline__ int xcd_remap_a(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

template <int MODE>
__global__ void __launch_bounds__(128) k_acoustic(AcousticArgs A, GLLk G) {
  __shared__ realw sC[128];               // potential tile
  __shared__ realw sT[3][128];            // products t1, t2, t3
  __shared__ realw sG[75];                // hprime, hprimewgll, wgllwgll
  const int p = threadIdx.x;
  const bool active = p < 125;
  if (p < 25) sG[p] = G.hp[p];
  else if (p < 50) sG[p] = G.hpw[p - 25];
  else if (p < 75) sG[p] = G.wxy[p - 50];
  const int ie = xcd_remap_a(blockIdx.x, gridDim.x);
  const int e = A.elems[ie];
  const int q = active ? p : 124;
  const int k = q / 25, j = (q / 5) % 5, i = q % 5;
  const int64_t ep = (int64_t)e * 125 + q;
  const int g = A.ibool[ep];
  const int irr = A.irregular[e];
  const realw rho_invl = (realw)1 / A.rho[ep];
  realw xixl = 0, xiyl = 0, xizl = 0, etaxl = 0, etayl = 0, etazl = 0, gaxl = 0, gayl = 0, gazl = 0, jl = 0;
  if (irr) {
    const int64_t qq = (int64_t)(irr - 1) * 125 + q;
    const int64_t P = A.jac_plane;
    const realw* J = A.jac;
    xixl = J[qq]; xiyl = J[P + qq]; xizl = J[2 * P + qq];
    etaxl = J[3 * P + qq]; etayl = J[4 * P + qq]; etazl = J[5 * P + qq];
    gaxl = J[6 * P + qq]; gayl = J[7 * P + qq]; gazl = J[8 * P + qq];
    jl = J[9 * P + qq];
  }
  sC[p] = A.pot[g];
  __syncthreads();
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
  realw a1 = sG[i * 5] * sC[b1];
  realw a2 = sG[j * 5] * sC[b2];
  realw a3 = sG[k * 5] * sC[b3];
#pragma unroll
  for (int l = 1; l < 5; l++) {
    a1 = a1 + sG[i * 5 + l] * sC[b1 + l];
    a2 = a2 + sG[j * 5 + l] * sC[b2 + 5 * l];
    a3 = a3 + sG[k * 5 + l] * sC[b3 + 25 * l];
  }
  realw t1, t2, t3;
  if (irr) {
    const realw dpdx = xixl * a1 + etaxl * a2 + gaxl * a3;
    const realw dpdy = xiyl * a1 + etayl * a2 + gayl * a3;
    const realw dpdz = xizl * a1 + etazl * a2 + gazl * a3;
    t1 = rho_invl * jl * (xixl * dpdx + xiyl * dpdy + xizl * dpdz);
    t2 = rho_invl * jl * (etaxl * dpdx + etayl * dpdy + etazl * dpdz);
    t3 = rho_invl * jl * (gaxl * dpdx + gayl * dpdy + gazl * dpdz);
  } else {
    const realw jr = A.jacobian_regular, xr = A.xix_regular;
    t1 = rho_invl * jr * xr * xr * a1;
    t2 = rho_invl * jr * xr * xr * a2;
    t3 = rho_invl * jr * xr * xr * a3;
  }
  sT[0][p] = t1;
  sT[1][p] = t2;
  sT[2][p] = t3;
  __syncthreads();
  if (!active) return;
  const realw* hpw = sG + 25;
  realw n1 = hpw[i] * sT[0][b1];
  realw n2 = hpw[j] * sT[1][b2];
  realw n3 = hpw[k] * sT[2][b3];
#pragma unroll
  for (int l = 1; l < 5; l++) {
    n1 = n1 + hpw[l * 5 + i] * sT[0][b1 + l];
    n2 = n2 + hpw[l * 5 + j] * sT[1][b2 + 5 * l];
    n3 = n3 + hpw[l * 5 + k] * sT[2][b3 + 25 * l];
  }
  const realw* wxy = sG + 50;
  const realw f = wxy[j * 5 + k] * n1 + wxy[i * 5 + k] * n2 + wxy[i * 5 + j] * n3;
  if (MODE == 1) {
    A.local_out[ep] = f;
  } else if (MODE == 2) {
    A.potdd[g] = A.potdd[g] + (-f);
  } else {
    atomicAdd(A.potdd + g, -f);
  }
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_acoustic)(hipStream_t s, const AcousticArgs& a, int mode) {
  if (a.nelem <= 0) return;
  dim3 grid(a.nelem), block(128);
  if (mode == 1)
    hipLaunchKernelGGL(k_acoustic<1>, grid, block, 0, s, a, g_gll_host);
  else if (mode == 2)
    hipLaunchKernelGGL(k_acoustic<2>, grid, block, 0, s, a, g_gll_host);
  else
    hipLaunchKernelGGL(k_acoustic<0>, grid, block, 0, s, a, g_gll_host);
  SF_LAUNCH_CHECK("k_acoustic");
}

}  // namespace sf
