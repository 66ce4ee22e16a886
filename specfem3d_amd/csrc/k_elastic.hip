// k_elastic.hip -- isotropic elastic / viscoelastic (3 SLS) element stiffness for NGLL=5.
//
// Restates compute_forces_viscoelastic.F90:33-962 (with the memory update of
// compute_element_att_memory.f90:28-156) as one gfx950 kernel:
//   * a 128-thread workgroup (two wave64s) processes a tile of 5 elements; lane t works on
//     element t/25 and on the GLL column (i, j) = (t%5, (t%25)/5), holding the 5 points of the
//     column (k = 0..4) in registers, so 125 of the 128 lanes are busy;
//   * the gamma-direction contractions run in registers; the xi/eta ones read the element
//     tile staged once in LDS (6 planes x 128 floats per element: u and dt*v, later the
//     stress terms t1/t2);
//   * the element-to-global scatter is selected by MODE: 0 a float atomic add, 1 a store of
//     the local contributions that a gather kernel later sums in the CPU's element order,
//     2 a plain read-modify-write (the launch holds one colour of a greedy element colouring,
//     so no two of its elements share a global point);
//   * workgroup ids are remapped so each XCD (8 per chip, dispatched round-robin) walks one
//     contiguous range of elements and keeps the shared points of neighbours in its own L2.
// The arithmetic order of every sum follows the reference (the five terms of each
// contraction left to right, etc.). This file is compiled twice: with -ffp-contract=off
// (SF_EXACT=1, same rounding as the CPU reference) and with FMA contraction (SF_EXACT=0).
#include "elastic_point.h"

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
// each build variant keeps its kernels in its own namespace (template instantiations of the
// two variants would otherwise be merged by the linker)
namespace SF_VNS {

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  // bijection: blocks b, b+8, b+16, ... (one XCD) get one contiguous logical range
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

template <bool STRAIN, bool ATT, int MODE>
__global__ void __launch_bounds__(128) k_elastic(ElasticArgs A, GLLk G) {
  __shared__ realw s_buf[5][6][128];
  const int tid = threadIdx.x;
  const int el = tid / 25;
  const int c = tid - el * 25;
  const int i = c % 5, j = c / 5;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ie = blk * 5 + el;
  const bool active = (tid < 125) && (ie < A.nelem);
  const int e = active ? A.elems[ie] : 0;
  const int elx = el < 5 ? el : 4;

  realw hpi[5], hpj[5], hpwi[5], hpwj[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpi[l] = G.hp[i * 5 + l];
    hpj[l] = G.hp[j * 5 + l];
    hpwi[l] = G.hpw[l * 5 + i];
    hpwj[l] = G.hpw[l * 5 + j];
  }

  int g[5];
  realw ux[5], uy[5], uz[5];
  realw vx[5], vy[5], vz[5];
  if (active) {
    const int* ib = A.ibool + (int64_t)e * 125 + c;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[k] = ib[25 * k];
      const realw* d = A.displ + 3 * (int64_t)g[k];
      ux[k] = d[0]; uy[k] = d[1]; uz[k] = d[2];
      s_buf[el][0][c + 25 * k] = ux[k];
      s_buf[el][1][c + 25 * k] = uy[k];
      s_buf[el][2][c + 25 * k] = uz[k];
      if (ATT) {
        const realw* v = A.veloc + 3 * (int64_t)g[k];
        vx[k] = A.deltat * v[0]; vy[k] = A.deltat * v[1]; vz[k] = A.deltat * v[2];
        s_buf[el][3][c + 25 * k] = vx[k];
        s_buf[el][4][c + 25 * k] = vy[k];
        s_buf[el][5][c + 25 * k] = vz[k];
      }
    }
  }
  __syncthreads();

  const int irr = active ? A.irregular[e] : 0;
  realw t1[3][5], t2[3][5], t3[3][5];
  const realw ONE_THIRD = (realw)1 / (realw)3;
  const realw FOUR_THIRDS = (realw)4 / (realw)3;

#pragma unroll
  for (int k = 0; k < 5; k++) {
    if (!active) continue;
    const int p = c + 25 * k;
    const realw* sx = s_buf[elx][0];
    const realw* sy = s_buf[elx][1];
    const realw* sz = s_buf[elx][2];
    const int b1 = 5 * j + 25 * k;  // f(l, j, k) = s[b1 + l]
    const int b2 = i + 25 * k;      // f(i, l, k) = s[b2 + 5 l]
    realw tx1 = hpi[0] * sx[b1], ty1 = hpi[0] * sy[b1], tz1 = hpi[0] * sz[b1];
    realw tx2 = hpj[0] * sx[b2], ty2 = hpj[0] * sy[b2], tz2 = hpj[0] * sz[b2];
    realw tx3 = G.hp[k * 5] * ux[0], ty3 = G.hp[k * 5] * uy[0], tz3 = G.hp[k * 5] * uz[0];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      tx1 = tx1 + hpi[l] * sx[b1 + l]; ty1 = ty1 + hpi[l] * sy[b1 + l]; tz1 = tz1 + hpi[l] * sz[b1 + l];
      tx2 = tx2 + hpj[l] * sx[b2 + 5 * l]; ty2 = ty2 + hpj[l] * sy[b2 + 5 * l]; tz2 = tz2 + hpj[l] * sz[b2 + 5 * l];
      tx3 = tx3 + G.hp[k * 5 + l] * ux[l]; ty3 = ty3 + G.hp[k * 5 + l] * uy[l]; tz3 = tz3 + G.hp[k * 5 + l] * uz[l];
    }

    realw xixl = 0, xiyl = 0, xizl = 0, etaxl = 0, etayl = 0, etazl = 0, gaxl = 0, gayl = 0, gazl = 0, jacl;
    realw duxdx, duxdy, duxdz, duydx, duydy, duydz, duzdx, duzdy, duzdz;
    if (irr) {
      const int64_t q = (int64_t)(irr - 1) * 125 + p;
      const realw* J = A.jac;
      const int64_t P = A.jac_plane;
      xixl = J[q]; xiyl = J[P + q]; xizl = J[2 * P + q];
      etaxl = J[3 * P + q]; etayl = J[4 * P + q]; etazl = J[5 * P + q];
      gaxl = J[6 * P + q]; gayl = J[7 * P + q]; gazl = J[8 * P + q];
      jacl = J[9 * P + q];
      duxdx = xixl * tx1 + etaxl * tx2 + gaxl * tx3;
      duxdy = xiyl * tx1 + etayl * tx2 + gayl * tx3;
      duxdz = xizl * tx1 + etazl * tx2 + gazl * tx3;
      duydx = xixl * ty1 + etaxl * ty2 + gaxl * ty3;
      duydy = xiyl * ty1 + etayl * ty2 + gayl * ty3;
      duydz = xizl * ty1 + etazl * ty2 + gazl * ty3;
      duzdx = xixl * tz1 + etaxl * tz2 + gaxl * tz3;
      duzdy = xiyl * tz1 + etayl * tz2 + gayl * tz3;
      duzdz = xizl * tz1 + etazl * tz2 + gazl * tz3;
    } else {
      const realw xr = A.xix_regular;
      jacl = A.jacobian_regular;
      duxdx = xr * tx1; duxdy = xr * tx2; duxdz = xr * tx3;
      duydx = xr * ty1; duydy = xr * ty2; duydz = xr * ty3;
      duzdx = xr * tz1; duzdy = xr * tz2; duzdz = xr * tz3;
    }

    const int64_t ep = (int64_t)e * 125 + p;
    realw e_xx = 0, e_yy = 0, e_xy = 0, e_xz = 0, e_yz = 0, e_tr = 0;
    if (STRAIN) {
      realw sxx, sxy, sxz, syx, syy, syz, szx, szy, szz;
      if (ATT) {
        // strain at u + dt v: continue each contraction with the dt*v tile
        const realw* wx = s_buf[elx][3];
        const realw* wy = s_buf[elx][4];
        const realw* wz = s_buf[elx][5];
        realw a1 = tx1, a2 = tx2, a3 = tx3, c1 = ty1, c2 = ty2, c3 = ty3, d1 = tz1, d2 = tz2, d3 = tz3;
#pragma unroll
        for (int l = 0; l < 5; l++) {
          const realw h1 = hpi[l], h2 = hpj[l], h3 = G.hp[k * 5 + l];
          a1 = a1 + wx[b1 + l] * h1; c1 = c1 + wy[b1 + l] * h1; d1 = d1 + wz[b1 + l] * h1;
          a2 = a2 + wx[b2 + 5 * l] * h2; c2 = c2 + wy[b2 + 5 * l] * h2; d2 = d2 + wz[b2 + 5 * l] * h2;
          a3 = a3 + vx[l] * h3; c3 = c3 + vy[l] * h3; d3 = d3 + vz[l] * h3;
        }
        if (irr) {
          sxx = xixl * a1 + etaxl * a2 + gaxl * a3;
          sxy = xiyl * a1 + etayl * a2 + gayl * a3;
          sxz = xizl * a1 + etazl * a2 + gazl * a3;
          syx = xixl * c1 + etaxl * c2 + gaxl * c3;
          syy = xiyl * c1 + etayl * c2 + gayl * c3;
          syz = xizl * c1 + etazl * c2 + gazl * c3;
          szx = xixl * d1 + etaxl * d2 + gaxl * d3;
          szy = xiyl * d1 + etayl * d2 + gayl * d3;
          szz = xizl * d1 + etazl * d2 + gazl * d3;
        } else {
          const realw xr = A.xix_regular;
          sxx = xr * a1; sxy = xr * a2; sxz = xr * a3;
          syx = xr * c1; syy = xr * c2; syz = xr * c3;
          szx = xr * d1; szy = xr * d2; szz = xr * d3;
        }
      } else {
        sxx = duxdx; sxy = duxdy; sxz = duxdz; syx = duydx; syy = duydy; syz = duydz;
        szx = duzdx; szy = duzdy; szz = duzdz;
      }
      e_tr = (sxx + syy + szz);
      const realw templ = ONE_THIRD * (sxx + syy + szz);
      if (A.eps_tr3) A.eps_tr3[ep] = templ;
      e_xx = sxx - templ;
      e_yy = syy - templ;
      e_xy = (realw)0.5 * (sxy + syx);
      e_xz = (realw)0.5 * (szx + sxz);
      e_yz = (realw)0.5 * (szy + syz);
    }

    const float2 km = A.kappamu[ep];
    const realw kappal = km.x, mul = km.y;
    const realw lp2m = kappal + FOUR_THIRDS * mul;
    const realw lambdal = lp2m - (realw)2 * mul;
    realw sigma_xx = lp2m * duxdx + lambdal * (duydy + duzdz);
    realw sigma_yy = lp2m * duydy + lambdal * (duxdx + duzdz);
    realw sigma_zz = lp2m * duzdz + lambdal * (duxdx + duydy);
    realw sigma_xy = mul * (duxdy + duydx);
    realw sigma_xz = mul * (duzdx + duxdz);
    realw sigma_yz = mul * (duzdy + duydz);

    if (ATT) {
      const int64_t nl = A.nloc;
      realw* R = A.R;
      realw Rv[6][3];
#pragma unroll
      for (int cc = 0; cc < 6; cc++)
#pragma unroll
        for (int s = 0; s < 3; s++) Rv[cc][s] = R[r_idx(e, cc * 3 + s, p)];
      const realw rxx = (Rv[0][0] + Rv[0][1]) + Rv[0][2];
      const realw ryy = (Rv[1][0] + Rv[1][1]) + Rv[1][2];
      const realw rtr = (Rv[5][0] + Rv[5][1]) + Rv[5][2];
      sigma_xx = sigma_xx - rxx - rtr;
      sigma_yy = sigma_yy - ryy - rtr;
      sigma_zz = sigma_zz + rxx + ryy - rtr;
      sigma_xy = sigma_xy - ((Rv[2][0] + Rv[2][1]) + Rv[2][2]);
      sigma_xz = sigma_xz - ((Rv[3][0] + Rv[3][1]) + Rv[3][2]);
      sigma_yz = sigma_yz - ((Rv[4][0] + Rv[4][1]) + Rv[4][2]);
      // memory-variable update (second-order RK) with old strain Sn and new strain Snp1
      realw* Ep = A.eps;
      const realw Sn[6] = {Ep[eps_idx(e, 0, p)], Ep[eps_idx(e, 1, p)], Ep[eps_idx(e, 2, p)], Ep[eps_idx(e, 3, p)],
                           Ep[eps_idx(e, 4, p)], Ep[eps_idx(e, 5, p)]};
      const realw Snp1[6] = {e_xx, e_yy, e_xy, e_xz, e_yz, e_tr};
#pragma unroll
      for (int s = 0; s < 3; s++) {
        const realw flk = kappal * A.fck[s * nl + ep];
        R[r_idx(e, 5 * 3 + s, p)] = A.alpha[s] * Rv[5][s] + flk * (A.beta[s] * Sn[5] + A.gamma[s] * Snp1[5]);
        const realw flm = mul * A.fc[s * nl + ep];
#pragma unroll
        for (int cc = 0; cc < 5; cc++)
          R[r_idx(e, cc * 3 + s, p)] = A.alpha[s] * Rv[cc][s] + flm * (A.beta[s] * Sn[cc] + A.gamma[s] * Snp1[cc]);
      }
    }
    if (STRAIN) {
      realw* Ep = A.eps;
      Ep[eps_idx(e, 0, p)] = e_xx; Ep[eps_idx(e, 1, p)] = e_yy; Ep[eps_idx(e, 2, p)] = e_xy;
      Ep[eps_idx(e, 3, p)] = e_xz; Ep[eps_idx(e, 4, p)] = e_yz;
      if (ATT) Ep[eps_idx(e, 5, p)] = e_tr;
    }

    const realw syx = sigma_xy, szx = sigma_xz, szy = sigma_yz;
    if (irr) {
      t1[0][k] = jacl * (sigma_xx * xixl + syx * xiyl + szx * xizl);
      t1[1][k] = jacl * (sigma_xy * xixl + sigma_yy * xiyl + szy * xizl);
      t1[2][k] = jacl * (sigma_xz * xixl + sigma_yz * xiyl + sigma_zz * xizl);
      t2[0][k] = jacl * (sigma_xx * etaxl + syx * etayl + szx * etazl);
      t2[1][k] = jacl * (sigma_xy * etaxl + sigma_yy * etayl + szy * etazl);
      t2[2][k] = jacl * (sigma_xz * etaxl + sigma_yz * etayl + sigma_zz * etazl);
      t3[0][k] = jacl * (sigma_xx * gaxl + syx * gayl + szx * gazl);
      t3[1][k] = jacl * (sigma_xy * gaxl + sigma_yy * gayl + szy * gazl);
      t3[2][k] = jacl * (sigma_xz * gaxl + sigma_yz * gayl + sigma_zz * gazl);
    } else {
      const realw xr = A.xix_regular;
      t1[0][k] = jacl * sigma_xx * xr;
      t1[1][k] = jacl * sigma_xy * xr;
      t1[2][k] = jacl * sigma_xz * xr;
      t2[0][k] = jacl * syx * xr;
      t2[1][k] = jacl * sigma_yy * xr;
      t2[2][k] = jacl * sigma_yz * xr;
      t3[0][k] = jacl * szx * xr;
      t3[1][k] = jacl * szy * xr;
      t3[2][k] = jacl * sigma_zz * xr;
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        s_buf[el][cc][c + 25 * k] = t1[cc][k];
        s_buf[el][3 + cc][c + 25 * k] = t2[cc][k];
      }
  }
  __syncthreads();
  if (!active) return;

#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int b1 = 5 * j + 25 * k;
    const int b2 = i + 25 * k;
    realw f[3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      const realw* s1 = s_buf[el][cc];
      const realw* s2 = s_buf[el][3 + cc];
      realw n1 = hpwi[0] * s1[b1];
      realw n2 = hpwj[0] * s2[b2];
      realw n3 = G.hpw[k] * t3[cc][0];
#pragma unroll
      for (int l = 1; l < 5; l++) {
        n1 = n1 + hpwi[l] * s1[b1 + l];
        n2 = n2 + hpwj[l] * s2[b2 + 5 * l];
        n3 = n3 + G.hpw[l * 5 + k] * t3[cc][l];
      }
      const realw fac1 = G.wxy[j * 5 + k], fac2 = G.wxy[i * 5 + k], fac3 = G.wxy[i * 5 + j];
      f[cc] = fac1 * n1 + fac2 * n2 + fac3 * n3;
    }
    if (MODE == 1) {
      const int64_t ep = (int64_t)e * 125 + c + 25 * k;
      A.local_out[ep] = f[0];
      A.local_out[A.nloc + ep] = f[1];
      A.local_out[2 * A.nloc + ep] = f[2];
    } else if (MODE == 2) {
      realw* a = A.accel + 3 * (int64_t)g[k];
      a[0] = a[0] + (-f[0]);
      a[1] = a[1] + (-f[1]);
      a[2] = a[2] + (-f[2]);
    } else {
      realw* a = A.accel + 3 * (int64_t)g[k];
      atomicAdd(a + 0, -f[0]);
      atomicAdd(a + 1, -f[1]);
      atomicAdd(a + 2, -f[2]);
    }
  }
}

// stored strain (6 planes) and SLS factors (table row or planes) of one point
__device__ __forceinline__ void load_factors(const ElasticArgs& A, int64_t ep, int64_t nl, realw fcv[3],
                                             realw fckv[3]) {
  if (A.fc_index) {
    const int r = A.fc_index[ep];
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = A.fc_table[s * A.fc_rows + r];
      fckv[s] = A.fc_table[(3 + s) * A.fc_rows + r];
    }
  } else {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = A.fc[s * nl + ep];
      fckv[s] = A.fck[s * nl + ep];
    }
  }
}

__device__ __forceinline__ void load_att_state(const ElasticArgs& A, int e, int q, int64_t ep, int64_t nl, realw Sn[6],
                                               realw fcv[3], realw fckv[3]) {
#pragma unroll
  for (int cc = 0; cc < 6; cc++) Sn[cc] = A.eps[eps_idx(e, cc, q)];
  if (A.fc_index) {
    const int r = A.fc_index[ep];
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = A.fc_table[s * A.fc_rows + r];
      fckv[s] = A.fc_table[(3 + s) * A.fc_rows + r];
    }
  } else {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = A.fc[s * nl + ep];
      fckv[s] = A.fck[s * nl + ep];
    }
  }
}

// Point-per-lane variant: a 128-thread workgroup (two wave64s) processes one element and lane p
// holds the GLL point p = i + 5 j + 25 k (lanes 125..127 shadow point 124 and store nothing).
// All three contractions read the element tile from LDS, so no lane carries a column of state
// between phases; the per-point state (memory variables, stored strain, SLS factors, moduli,
// Jacobian terms) is loaded before the first barrier and is in flight while the tile is
// gathered. Same arithmetic, in the same order, as k_elastic.
// AV selects how the attenuation state is scheduled (the "att_variant" tuning switch; same
// results; schedule 5 is the persistent kernel k_elastic_pipe below):
// 0 every per-point array is loaded before the first barrier; 1 only the memory variables are,
// the stored strain and SLS factors are loaded where the memory update needs them (fewer live
// registers); 2 as 1 with the register budget capped for 5 waves per SIMD; 3 as 0 with that cap;
// 4 loads the memory variables and stored strain straight into LDS (global_load_lds, no
// registers held while they are in flight; each lane reads back its own point), which frees
// registers for 5 waves per SIMD; the stress products then reuse those LDS planes.
typedef __attribute__((address_space(3))) void* sf_lds_ptr;
typedef __attribute__((address_space(1))) void* sf_global_ptr;

// One dword per lane from global memory straight into LDS at lds_base + 4 * lane, issued as
// assembly so the compiler does not track it: its waits for LDS-direct data would be vmcnt(0)
// before every read of any LDS array it cannot tell apart from the target. The caller waits
// (vmcnt) before the lane reads its slot back. (M0 needs one wait state before the load.)
__device__ __forceinline__ void lds_dma_dword(const realw* g, uint32_t lds_base) {
  const uint32_t b = __builtin_amdgcn_readfirstlane(lds_base);  // wave-uniform
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off" ::"v"(g), "s"(b) : "memory", "m0");
}

template <bool STRAIN, bool ATT, int MODE, int AV>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(AV >= 2 ? 5 : 1)))
k_elastic_pt(ElasticArgs A, GLLk G) {
  constexpr bool LATE = ATT && (AV == 1 || AV == 2);
  constexpr bool DL = ATT && AV == 4;
  __shared__ realw sU[6][128];            // u (0..2) and dt*v (3..5)
  __shared__ realw sX[DL ? 24 : 9][128];  // stress products (planes 0..8); with DL first the state
  realw(*const sT)[128] = sX;             // t1 (0..2), t2 (3..5), t3 (6..8)
  __shared__ realw sG[75];                // hprime, hprimewgll, wgllwgll
  const int p = threadIdx.x;
  const bool active = p < 125;
  stage_gll(sG, G, p);
  const int ie = xcd_remap(blockIdx.x, gridDim.x);  // grid = one workgroup per listed element
  const int e = A.elems[ie];
  const int q = active ? p : 124;
  const int k = q / 25, j = (q / 5) % 5, i = q % 5;
  const int64_t ep = (int64_t)e * 125 + q;
  const int64_t nl = A.nloc;
  const int g = A.ibool[ep];
  const int irr = A.irregular[e];
  JacPt J = {};
  realw jacl = A.jacobian_regular;
  if (irr) {
    J = load_jac(A.jac, A.jac_plane, (int64_t)(irr - 1) * 125 + q);
    jacl = J.jac;
  }
  const float2 km = A.kappamu[ep];
  realw Rv[6][3], Sn[6], fcv[3], fckv[3];
  if (ATT && !DL) {
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int s = 0; s < 3; s++) Rv[cc][s] = A.R[r_idx(e, cc * 3 + s, q)];
    if (!LATE) load_att_state(A, e, q, ep, nl, Sn, fcv, fckv);
  }
  if (DL) load_factors(A, ep, nl, fcv, fckv);
  {
    const realw* d = A.displ + 3 * (int64_t)g;
    const realw u0 = d[0], u1 = d[1], u2 = d[2];
    realw v0 = 0, v1 = 0, v2 = 0;
    if (ATT) {
      const realw* v = A.veloc + 3 * (int64_t)g;
      v0 = v[0]; v1 = v[1]; v2 = v[2];
    }
    sU[0][p] = u0; sU[1][p] = u1; sU[2][p] = u2;
    if (ATT) {
      sU[3][p] = A.deltat * v0; sU[4][p] = A.deltat * v1; sU[5][p] = A.deltat * v2;
    }
  }
  __syncthreads();
  if (DL) {
    // after the barrier (whose release would wait for them, as they write LDS): R planes 0..17 and
    // strain planes 0..5 of this element into sX[plane][lane], in flight during the gradient;
    // each lane reads back only its own slots, after its own wait
    __builtin_amdgcn_sched_barrier(0);
    const int w64 = p & 64;
#pragma unroll
    for (int pl = 0; pl < 18; pl++)
      __builtin_amdgcn_global_load_lds((sf_global_ptr)(A.R + r_idx(e, pl, q)), (sf_lds_ptr)&sX[pl][w64], 4, 0, 0);
#pragma unroll
    for (int c = 0; c < 6; c++)
      __builtin_amdgcn_global_load_lds((sf_global_ptr)(A.eps + eps_idx(e, c, q)), (sf_lds_ptr)&sX[18 + c][w64], 4,
                                       0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }

  const realw ONE_THIRD = (realw)1 / (realw)3;
  const Grad9 tg = tile_gradient<1>(sU[0], sU[1], sU[2], sG, i, j, k);
  const Du9 du = map_gradient(tg, irr != 0, J, A.xix_regular);

  realw e_xx = 0, e_yy = 0, e_xy = 0, e_xz = 0, e_yz = 0, e_tr = 0;
  if (STRAIN) {
    Du9 st = du;
    if (ATT) st = map_gradient(tile_gradient_continue<1>(tg, sU[3], sU[4], sU[5], sG, i, j, k), irr != 0, J,
                               A.xix_regular);
    e_tr = (st.xx + st.yy + st.zz);
    const realw templ = ONE_THIRD * (st.xx + st.yy + st.zz);
    if (A.eps_tr3 && active) A.eps_tr3[ep] = templ;
    e_xx = st.xx - templ;
    e_yy = st.yy - templ;
    e_xy = (realw)0.5 * (st.xy + st.yx);
    e_xz = (realw)0.5 * (st.zx + st.xz);
    e_yz = (realw)0.5 * (st.zy + st.yz);
  }

  const realw kappal = km.x, mul = km.y;
  Sigma6 sg = iso_stress(du, kappal, mul);

  if (DL) {
    __builtin_amdgcn_s_waitcnt(0);  // the LDS-direct loads of this wave have landed
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int s = 0; s < 3; s++) Rv[cc][s] = sX[cc * 3 + s][p];
#pragma unroll
    for (int c = 0; c < 6; c++) Sn[c] = sX[18 + c][p];
  }
  if (ATT) {
    const realw rxx = (Rv[0][0] + Rv[0][1]) + Rv[0][2];
    const realw ryy = (Rv[1][0] + Rv[1][1]) + Rv[1][2];
    const realw rtr = (Rv[5][0] + Rv[5][1]) + Rv[5][2];
    sg.xx = sg.xx - rxx - rtr;
    sg.yy = sg.yy - ryy - rtr;
    sg.zz = sg.zz + rxx + ryy - rtr;
    sg.xy = sg.xy - ((Rv[2][0] + Rv[2][1]) + Rv[2][2]);
    sg.xz = sg.xz - ((Rv[3][0] + Rv[3][1]) + Rv[3][2]);
    sg.yz = sg.yz - ((Rv[4][0] + Rv[4][1]) + Rv[4][2]);
    if (active) {
      // memory-variable update (second-order RK) with old strain Sn and new strain Snp1
      if (LATE) load_att_state(A, e, q, ep, nl, Sn, fcv, fckv);
      const realw Snp1[6] = {e_xx, e_yy, e_xy, e_xz, e_yz, e_tr};
#pragma unroll
      for (int s = 0; s < 3; s++) {
        const realw flk = kappal * fckv[s];
        A.R[r_idx(e, 5 * 3 + s, q)] = A.alpha[s] * Rv[5][s] + flk * (A.beta[s] * Sn[5] + A.gamma[s] * Snp1[5]);
        const realw flm = mul * fcv[s];
#pragma unroll
        for (int cc = 0; cc < 5; cc++)
          A.R[r_idx(e, cc * 3 + s, q)] = A.alpha[s] * Rv[cc][s] + flm * (A.beta[s] * Sn[cc] + A.gamma[s] * Snp1[cc]);
      }
    }
  }
  if (STRAIN && active) {
    realw* Ep = A.eps;
    Ep[eps_idx(e, 0, q)] = e_xx; Ep[eps_idx(e, 1, q)] = e_yy; Ep[eps_idx(e, 2, q)] = e_xy;
    Ep[eps_idx(e, 3, q)] = e_xz; Ep[eps_idx(e, 4, q)] = e_yz;
    if (ATT) Ep[eps_idx(e, 5, q)] = e_tr;
  }

  {
    realw t[9];
    stress_products(sg, irr != 0, J, jacl, A.xix_regular, t);
#pragma unroll
    for (int m = 0; m < 9; m++) sT[m][p] = t[m];
  }
  lds_barrier();  // the memory-variable and strain stores drain during the divergence
  if (!active) return;

  realw f[3];
  tile_divergence<1>(&sT[0][0], 128, sG + 25, sG + 50, i, j, k, f);
  if (MODE == 1) {
    A.local_out[ep] = f[0];
    A.local_out[nl + ep] = f[1];
    A.local_out[2 * nl + ep] = f[2];
  } else if (MODE == 2) {
    // coloured launch. (A plain update of the element-interior points inside atomic launches
    // measured 1.9x slower on the 96x96x48 viscoelastic workload than atomics for all points.)
    realw* a = A.accel + 3 * (int64_t)g;
    a[0] = a[0] + (-f[0]);
    a[1] = a[1] + (-f[1]);
    a[2] = a[2] + (-f[2]);
  } else {
    realw* a = A.accel + 3 * (int64_t)g;
    atomicAdd(a + 0, -f[0]);
    atomicAdd(a + 1, -f[1]);
    atomicAdd(a + 2, -f[2]);
  }
}

// Schedule 5 (viscoelastic only): a persistent workgroup walks the elements ie, ie + NG, ie + 2 NG,
// ... (NG = grid size) as a software pipeline, so that the next element's loads are in flight
// while this element is computed: its displacement and velocity, moduli and factor row are
// loaded into registers at the top of the iteration, and its memory variables and stored strain
// go straight into LDS (global_load_lds) once this lane has read back its own. Element indices
// run three deep (the list entry two elements ahead, the ibool row one ahead) so no load in the
// loop waits for another. The stress products have LDS planes of their own, so the state planes
// are free for the next element as soon as they are read. Barriers order LDS only (lds_barrier),
// leaving global loads, stores and atomics in flight. Per element the arithmetic and its order
// are k_elastic_pt's, so the results are the same bits.
template <int MODE>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4)))
k_elastic_pipe(ElasticArgs A, GLLk G) {
  __shared__ realw sU[6][128];   // u (0..2) and dt*v (3..5) of the current element
  __shared__ realw sX[24][128];  // memory variables (0..17) and stored strain (18..23), by DMA
  __shared__ realw sT[9][128];   // stress products
  __shared__ realw sG[75];
  const int p = threadIdx.x;
  const bool active = p < 125;
  stage_gll(sG, G, p);
  const int q = active ? p : 124;
  const int k = q / 25, j = (q / 5) % 5, i = q % 5;
  const int w64 = p & 64;
  const int64_t nl = A.nloc;
  const int NG = gridDim.x;
  const int nelem = A.nelem;
  int ie = xcd_remap(blockIdx.x, NG);
  if (ie >= nelem) return;  // the launcher keeps NG <= nelem
  const bool tab = A.fc_index != nullptr;
  // element indices of this and the next two iterations (a finished walk repeats its last
  // element in the look-ahead slots, so every load stays in bounds; nothing is written for them)
  int ie_v = ie;
  asm volatile("" : "+v"(ie_v));
  int e = A.elems[ie_v];
  int e_n = ie + NG < nelem ? A.elems[ie_v + NG] : e;
  int e_nn = ie + 2 * NG < nelem ? A.elems[ie_v + 2 * NG] : e_n;
  int64_t ep = (int64_t)e * 125 + q;
  int g = A.ibool[ep];
  int g_n = A.ibool[(int64_t)e_n * 125 + q];
  {
    const realw* d = A.displ + 3 * (int64_t)g;
    const realw* v = A.veloc + 3 * (int64_t)g;
    const realw u0 = d[0], u1 = d[1], u2 = d[2], v0 = v[0], v1 = v[1], v2 = v[2];
    sU[0][p] = u0; sU[1][p] = u1; sU[2][p] = u2;
    sU[3][p] = A.deltat * v0; sU[4][p] = A.deltat * v1; sU[5][p] = A.deltat * v2;
  }
  float2 km = A.kappamu[ep];
  int irr = A.irregular[e];
  int fcr = tab ? (int)A.fc_index[ep] : 0;
  asm volatile("" : "+v"(irr));
  const uint32_t xb = (uint32_t)(uintptr_t)(sf_lds_ptr)&sX[0][w64];  // this wave's slots of plane 0
#pragma unroll
  for (int pl = 0; pl < 18; pl++) lds_dma_dword(A.R + r_idx(e, pl, q), xb + 512 * pl);
#pragma unroll
  for (int c = 0; c < 6; c++) lds_dma_dword(A.eps + eps_idx(e, c, q), xb + 512 * (18 + c));
  lds_barrier();

  const realw ONE_THIRD = (realw)1 / (realw)3;
  for (bool first = true;; first = false) {
    const bool has_next = ie + NG < nelem;  // uniform over the workgroup
    // next element: u, v, moduli, factor row, Jacobian index; the ibool row after it; the list
    // entry after that. This element: its factor-table row and Jacobian terms.
    const int64_t ep_n = (int64_t)e_n * 125 + q;
    const int g_nn = A.ibool[(int64_t)e_nn * 125 + q];
    // list entries and Jacobian indices are read through a vector register: loaded from a
    // uniform address they would be moved to a scalar register right after the load, which
    // waits for every memory operation issued before it (this element's state, the previous
    // element's stores and atomics)
    int ie3 = ie + 3 * NG < nelem ? ie + 3 * NG : ie;
    asm volatile("" : "+v"(ie3));
    const int e_nnn = ie + 3 * NG < nelem ? A.elems[ie3] : e_nn;
    realw un0, un1, un2, vn0, vn1, vn2;
    {
      const realw* d = A.displ + 3 * (int64_t)g_n;
      const realw* v = A.veloc + 3 * (int64_t)g_n;
      un0 = d[0]; un1 = d[1]; un2 = d[2]; vn0 = v[0]; vn1 = v[1]; vn2 = v[2];
    }
    const float2 km_n = A.kappamu[ep_n];
    int en_v = e_n;
    asm volatile("" : "+v"(en_v));
    const int irr_n = A.irregular[en_v];
    const int fcr_n = tab ? (int)A.fc_index[ep_n] : 0;
    realw fcv[3], fckv[3];
    if (tab) {
#pragma unroll
      for (int sl = 0; sl < 3; sl++) {
        fcv[sl] = A.fc_table[sl * A.fc_rows + fcr];
        fckv[sl] = A.fc_table[(3 + sl) * A.fc_rows + fcr];
      }
    } else {
#pragma unroll
      for (int sl = 0; sl < 3; sl++) {
        fcv[sl] = A.fc[sl * nl + ep];
        fckv[sl] = A.fck[sl * nl + ep];
      }
    }
    JacPt J = {};
    realw jacl = A.jacobian_regular;
    if (irr) {
      J = load_jac(A.jac, A.jac_plane, (int64_t)(irr - 1) * 125 + q);
      jacl = J.jac;
    }
    __builtin_amdgcn_sched_barrier(0);

    const Grad9 tg = tile_gradient<1>(sU[0], sU[1], sU[2], sG, i, j, k);
    const Du9 du = map_gradient(tg, irr != 0, J, A.xix_regular);
    const Du9 st = map_gradient(tile_gradient_continue<1>(tg, sU[3], sU[4], sU[5], sG, i, j, k), irr != 0, J,
                                A.xix_regular);
    const realw e_tr = (st.xx + st.yy + st.zz);
    const realw templ = ONE_THIRD * (st.xx + st.yy + st.zz);
    if (A.eps_tr3 && active) A.eps_tr3[ep] = templ;
    const realw Snp1[6] = {st.xx - templ, st.yy - templ, (realw)0.5 * (st.xy + st.yx), (realw)0.5 * (st.zx + st.xz),
                           (realw)0.5 * (st.zy + st.yz), e_tr};
    const realw kappal = km.x, mul = km.y;
    Sigma6 sg = iso_stress(du, kappal, mul);

    // this element's state has landed; read own slots. In the first iteration everything is
    // waited for; later, at least the 24 memory-variable and strain stores of the previous
    // element were issued after this element's LDS-direct loads, so vmcnt(24) suffices and those
    // stores (and the previous element's assembly) stay in flight.
    if (first) {
      __builtin_amdgcn_s_waitcnt(0);
    } else {
      __builtin_amdgcn_s_waitcnt(0x4f78);  // vmcnt(24)
    }
    // The 24 slots are read in one asm block that ends with lgkmcnt(0): read through the
    // compiler, each read of a plane that an LDS-direct load wrote would get a vmcnt(0) first,
    // which would also wait for the previous element's stores and atomics.
    realw Rv[6][3], Sn[6];
    {
      const uint32_t sa = (uint32_t)(uintptr_t)(sf_lds_ptr)&sX[0][p];
      double x[12];
      asm volatile(
          "ds_read2st64_b32 %0, %12 offset0:0 offset1:2\n\t"
          "ds_read2st64_b32 %1, %12 offset0:4 offset1:6\n\t"
          "ds_read2st64_b32 %2, %12 offset0:8 offset1:10\n\t"
          "ds_read2st64_b32 %3, %12 offset0:12 offset1:14\n\t"
          "ds_read2st64_b32 %4, %12 offset0:16 offset1:18\n\t"
          "ds_read2st64_b32 %5, %12 offset0:20 offset1:22\n\t"
          "ds_read2st64_b32 %6, %12 offset0:24 offset1:26\n\t"
          "ds_read2st64_b32 %7, %12 offset0:28 offset1:30\n\t"
          "ds_read2st64_b32 %8, %12 offset0:32 offset1:34\n\t"
          "ds_read2st64_b32 %9, %12 offset0:36 offset1:38\n\t"
          "ds_read2st64_b32 %10, %12 offset0:40 offset1:42\n\t"
          "ds_read2st64_b32 %11, %12 offset0:44 offset1:46\n\t"
          "s_waitcnt lgkmcnt(0)"
          : "=v"(x[0]), "=v"(x[1]), "=v"(x[2]), "=v"(x[3]), "=v"(x[4]), "=v"(x[5]), "=v"(x[6]), "=v"(x[7]),
            "=v"(x[8]), "=v"(x[9]), "=v"(x[10]), "=v"(x[11])
          : "v"(sa)
          : "memory");
      realw pl[24];
#pragma unroll
      for (int h = 0; h < 12; h++) {
        const float2 f2 = __builtin_bit_cast(float2, x[h]);
        pl[2 * h] = f2.x;
        pl[2 * h + 1] = f2.y;
      }
#pragma unroll
      for (int cc = 0; cc < 6; cc++)
#pragma unroll
        for (int sl = 0; sl < 3; sl++) Rv[cc][sl] = pl[cc * 3 + sl];
#pragma unroll
      for (int c = 0; c < 6; c++) Sn[c] = pl[18 + c];
    }
    // every register operand of the memory update is taken before the next element's loads are
    // issued, so no wait after them can include those loads
    realw flk[3], flm[3];
#pragma unroll
    for (int sl = 0; sl < 3; sl++) {
      flk[sl] = km.x * fckv[sl];
      flm[sl] = km.y * fcv[sl];
    }
    asm volatile("" ::"v"(flk[0]), "v"(flk[1]), "v"(flk[2]), "v"(flm[0]), "v"(flm[1]), "v"(flm[2]));
    if (has_next) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int pl = 0; pl < 18; pl++) lds_dma_dword(A.R + r_idx(e_n, pl, q), xb + 512 * pl);
#pragma unroll
      for (int c = 0; c < 6; c++) lds_dma_dword(A.eps + eps_idx(e_n, c, q), xb + 512 * (18 + c));
      __builtin_amdgcn_sched_barrier(0);
    }

    const realw rxx = (Rv[0][0] + Rv[0][1]) + Rv[0][2];
    const realw ryy = (Rv[1][0] + Rv[1][1]) + Rv[1][2];
    const realw rtr = (Rv[5][0] + Rv[5][1]) + Rv[5][2];
    sg.xx = sg.xx - rxx - rtr;
    sg.yy = sg.yy - ryy - rtr;
    sg.zz = sg.zz + rxx + ryy - rtr;
    sg.xy = sg.xy - ((Rv[2][0] + Rv[2][1]) + Rv[2][2]);
    sg.xz = sg.xz - ((Rv[3][0] + Rv[3][1]) + Rv[3][2]);
    sg.yz = sg.yz - ((Rv[4][0] + Rv[4][1]) + Rv[4][2]);
    if (active) {
#pragma unroll
      for (int sl = 0; sl < 3; sl++) {
        A.R[r_idx(e, 5 * 3 + sl, q)] =
            A.alpha[sl] * Rv[5][sl] + flk[sl] * (A.beta[sl] * Sn[5] + A.gamma[sl] * Snp1[5]);
#pragma unroll
        for (int cc = 0; cc < 5; cc++)
          A.R[r_idx(e, cc * 3 + sl, q)] =
              A.alpha[sl] * Rv[cc][sl] + flm[sl] * (A.beta[sl] * Sn[cc] + A.gamma[sl] * Snp1[cc]);
      }
      realw* Ep = A.eps;
#pragma unroll
      for (int c = 0; c < 6; c++) Ep[eps_idx(e, c, q)] = Snp1[c];
    }
    {
      realw t[9];
      stress_products(sg, irr != 0, J, jacl, A.xix_regular, t);
#pragma unroll
      for (int m = 0; m < 9; m++) sT[m][p] = t[m];
    }
    lds_barrier();  // products complete; every lane is past this element's u tile
    if (has_next) {
      sU[0][p] = un0; sU[1][p] = un1; sU[2][p] = un2;
      sU[3][p] = A.deltat * vn0; sU[4][p] = A.deltat * vn1; sU[5][p] = A.deltat * vn2;
    }
    if (active) {
      realw f[3];
      tile_divergence<1>(&sT[0][0], 128, sG + 25, sG + 50, i, j, k, f);
      if (MODE == 1) {
        A.local_out[ep] = f[0];
        A.local_out[nl + ep] = f[1];
        A.local_out[2 * nl + ep] = f[2];
      } else if (MODE == 2) {
        realw* a = A.accel + 3 * (int64_t)g;
        a[0] = a[0] + (-f[0]);
        a[1] = a[1] + (-f[1]);
        a[2] = a[2] + (-f[2]);
      } else {
        realw* a = A.accel + 3 * (int64_t)g;
        atomicAdd(a + 0, -f[0]);
        atomicAdd(a + 1, -f[1]);
        atomicAdd(a + 2, -f[2]);
      }
    }
    lds_barrier();  // products read; the next element's tile is staged
    if (!has_next) break;
    ie += NG;
    e = e_n;
    e_n = e_nn;
    e_nn = e_nnn;
    ep = ep_n;
    g = g_n;
    g_n = g_nn;
    km = km_n;
    irr = irr_n;
    fcr = fcr_n;
  }
}

// Elastic (no attenuation) counterpart of k_elastic_pipe ("elastic_shape" 2): persistent
// workgroups walk ie, ie + NG, ... and load the next element's displacement, moduli and Jacobian
// index (and the ibool row of the one after it, and the list entry after that) while computing
// this one. Without memory variables the point kernel moves few bytes per element and waits on the
// dependent ibool -> displacement loads at the start of every element; here they are a whole
// element ahead. Per element the arithmetic and its order are k_elastic_pt's (same bits).
template <bool STRAIN, int MODE>
__global__ void __launch_bounds__(128) k_elastic_pipe_el(ElasticArgs A, GLLk G) {
  __shared__ realw sU[3][128];   // u of the current element
  __shared__ realw sT[9][128];   // stress products
  __shared__ realw sG[75];
  const int p = threadIdx.x;
  const bool active = p < 125;
  stage_gll(sG, G, p);
  const int q = active ? p : 124;
  const int k = q / 25, j = (q / 5) % 5, i = q % 5;
  const int64_t nl = A.nloc;
  const int NG = gridDim.x;
  const int nelem = A.nelem;
  int ie = xcd_remap(blockIdx.x, NG);
  if (ie >= nelem) return;  // the launcher keeps NG <= nelem
  int ie_v = ie;
  asm volatile("" : "+v"(ie_v));
  int e = A.elems[ie_v];
  int e_n = ie + NG < nelem ? A.elems[ie_v + NG] : e;
  int e_nn = ie + 2 * NG < nelem ? A.elems[ie_v + 2 * NG] : e_n;
  int64_t ep = (int64_t)e * 125 + q;
  int g = A.ibool[ep];
  int g_n = A.ibool[(int64_t)e_n * 125 + q];
  {
    const realw* d = A.displ + 3 * (int64_t)g;
    const realw u0 = d[0], u1 = d[1], u2 = d[2];
    sU[0][p] = u0; sU[1][p] = u1; sU[2][p] = u2;
  }
  float2 km = A.kappamu[ep];
  int irr = A.irregular[e];
  asm volatile("" : "+v"(irr));
  lds_barrier();

  const realw ONE_THIRD = (realw)1 / (realw)3;
  for (;;) {
    const bool has_next = ie + NG < nelem;  // uniform over the workgroup
    const int64_t ep_n = (int64_t)e_n * 125 + q;
    const int g_nn = A.ibool[(int64_t)e_nn * 125 + q];
    int ie3 = ie + 3 * NG < nelem ? ie + 3 * NG : ie;
    asm volatile("" : "+v"(ie3));
    const int e_nnn = ie + 3 * NG < nelem ? A.elems[ie3] : e_nn;
    realw un0, un1, un2;
    {
      const realw* d = A.displ + 3 * (int64_t)g_n;
      un0 = d[0]; un1 = d[1]; un2 = d[2];
    }
    const float2 km_n = A.kappamu[ep_n];
    int en_v = e_n;
    asm volatile("" : "+v"(en_v));
    const int irr_n = A.irregular[en_v];
    JacPt J = {};
    realw jacl = A.jacobian_regular;
    if (irr) {
      J = load_jac(A.jac, A.jac_plane, (int64_t)(irr - 1) * 125 + q);
      jacl = J.jac;
    }
    __builtin_amdgcn_sched_barrier(0);

    const Grad9 tg = tile_gradient<1>(sU[0], sU[1], sU[2], sG, i, j, k);
    const Du9 du = map_gradient(tg, irr != 0, J, A.xix_regular);
    if (STRAIN) {
      const Du9& st = du;
      const realw templ = ONE_THIRD * (st.xx + st.yy + st.zz);
      if (A.eps_tr3 && active) A.eps_tr3[ep] = templ;
      if (active) {
        realw* Ep = A.eps;
        Ep[eps_idx(e, 0, q)] = st.xx - templ;
        Ep[eps_idx(e, 1, q)] = st.yy - templ;
        Ep[eps_idx(e, 2, q)] = (realw)0.5 * (st.xy + st.yx);
        Ep[eps_idx(e, 3, q)] = (realw)0.5 * (st.zx + st.xz);
        Ep[eps_idx(e, 4, q)] = (realw)0.5 * (st.zy + st.yz);
      }
    }
    const Sigma6 sg = iso_stress(du, km.x, km.y);
    {
      realw t[9];
      stress_products(sg, irr != 0, J, jacl, A.xix_regular, t);
#pragma unroll
      for (int m = 0; m < 9; m++) sT[m][p] = t[m];
    }
    lds_barrier();  // products complete; every lane is past this element's u tile
    if (has_next) {
      sU[0][p] = un0; sU[1][p] = un1; sU[2][p] = un2;
    }
    if (active) {
      realw f[3];
      tile_divergence<1>(&sT[0][0], 128, sG + 25, sG + 50, i, j, k, f);
      if (MODE == 1) {
        A.local_out[ep] = f[0];
        A.local_out[nl + ep] = f[1];
        A.local_out[2 * nl + ep] = f[2];
      } else if (MODE == 2) {
        realw* a = A.accel + 3 * (int64_t)g;
        a[0] = a[0] + (-f[0]);
        a[1] = a[1] + (-f[1]);
        a[2] = a[2] + (-f[2]);
      } else {
        realw* a = A.accel + 3 * (int64_t)g;
        atomicAdd(a + 0, -f[0]);
        atomicAdd(a + 1, -f[1]);
        atomicAdd(a + 2, -f[2]);
      }
    }
    lds_barrier();  // products read; the next element's tile is staged
    if (!has_next) break;
    ie += NG;
    e = e_n;
    e_n = e_nn;
    e_nn = e_nnn;
    ep = ep_n;
    g = g_n;
    g_n = g_nn;
    km = km_n;
    irr = irr_n;
  }
}

// element strain of a displacement field (compute_element_strain.F90:28-130); one lane per point
__global__ void __launch_bounds__(128) k_element_strain(ElasticArgs A, GLLk G, realw* eps, realw* eps_tr3,
                                                        int with_trace) {
  __shared__ realw s_buf[5][3][128];
  const int tid = threadIdx.x;
  const int el = tid / 25;
  const int c = tid - el * 25;
  const int i = c % 5, j = c / 5;
  const int ie = blockIdx.x * 5 + el;
  const bool active = (tid < 125) && (ie < A.nelem);
  const int e = active ? A.elems[ie] : 0;
  realw ux[5], uy[5], uz[5];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int g = A.ibool[(int64_t)e * 125 + c + 25 * k];
      ux[k] = A.displ[3 * (int64_t)g]; uy[k] = A.displ[3 * (int64_t)g + 1]; uz[k] = A.displ[3 * (int64_t)g + 2];
      s_buf[el][0][c + 25 * k] = ux[k]; s_buf[el][1][c + 25 * k] = uy[k]; s_buf[el][2][c + 25 * k] = uz[k];
    }
  }
  __syncthreads();
  if (!active) return;
  const int irr = A.irregular[e];
  const realw ONE_THIRD = (realw)1 / (realw)3;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int p = c + 25 * k;
    const int b1 = 5 * j + 25 * k, b2 = i + 25 * k;
    const realw* sx = s_buf[el][0];
    const realw* sy = s_buf[el][1];
    const realw* sz = s_buf[el][2];
    realw tx1 = G.hp[i * 5] * sx[b1], ty1 = G.hp[i * 5] * sy[b1], tz1 = G.hp[i * 5] * sz[b1];
    realw tx2 = G.hp[j * 5] * sx[b2], ty2 = G.hp[j * 5] * sy[b2], tz2 = G.hp[j * 5] * sz[b2];
    realw tx3 = G.hp[k * 5] * ux[0], ty3 = G.hp[k * 5] * uy[0], tz3 = G.hp[k * 5] * uz[0];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      tx1 = tx1 + G.hp[i * 5 + l] * sx[b1 + l]; ty1 = ty1 + G.hp[i * 5 + l] * sy[b1 + l];
      tz1 = tz1 + G.hp[i * 5 + l] * sz[b1 + l];
      tx2 = tx2 + G.hp[j * 5 + l] * sx[b2 + 5 * l]; ty2 = ty2 + G.hp[j * 5 + l] * sy[b2 + 5 * l];
      tz2 = tz2 + G.hp[j * 5 + l] * sz[b2 + 5 * l];
      tx3 = tx3 + G.hp[k * 5 + l] * ux[l]; ty3 = ty3 + G.hp[k * 5 + l] * uy[l]; tz3 = tz3 + G.hp[k * 5 + l] * uz[l];
    }
    realw duxdx, duxdy, duxdz, duydx, duydy, duydz, duzdx, duzdy, duzdz;
    if (irr) {
      const int64_t q = (int64_t)(irr - 1) * 125 + p;
      const int64_t P = A.jac_plane;
      const realw* J = A.jac;
      const realw xixl = J[q], xiyl = J[P + q], xizl = J[2 * P + q], etaxl = J[3 * P + q], etayl = J[4 * P + q];
      const realw etazl = J[5 * P + q], gaxl = J[6 * P + q], gayl = J[7 * P + q], gazl = J[8 * P + q];
      duxdx = xixl * tx1 + etaxl * tx2 + gaxl * tx3;
      duxdy = xiyl * tx1 + etayl * tx2 + gayl * tx3;
      duxdz = xizl * tx1 + etazl * tx2 + gazl * tx3;
      duydx = xixl * ty1 + etaxl * ty2 + gaxl * ty3;
      duydy = xiyl * ty1 + etayl * ty2 + gayl * ty3;
      duydz = xizl * ty1 + etazl * ty2 + gazl * ty3;
      duzdx = xixl * tz1 + etaxl * tz2 + gaxl * tz3;
      duzdy = xiyl * tz1 + etayl * tz2 + gayl * tz3;
      duzdz = xizl * tz1 + etazl * tz2 + gazl * tz3;
    } else {
      const realw xr = A.xix_regular;
      duxdx = xr * tx1; duxdy = xr * tx2; duxdz = xr * tx3;
      duydx = xr * ty1; duydy = xr * ty2; duydz = xr * ty3;
      duzdx = xr * tz1; duzdy = xr * tz2; duzdz = xr * tz3;
    }
    const realw templ = ONE_THIRD * (duxdx + duydy + duzdz);
    const int64_t ep = (int64_t)e * 125 + p;
    eps_tr3[ep] = templ;
    eps[eps_idx(e, 0, p)] = duxdx - templ;
    eps[eps_idx(e, 1, p)] = duydy - templ;
    eps[eps_idx(e, 2, p)] = (realw)0.5 * (duxdy + duydx);
    eps[eps_idx(e, 3, p)] = (realw)0.5 * (duzdx + duxdz);
    eps[eps_idx(e, 4, p)] = (realw)0.5 * (duzdy + duydz);
    // with attenuation the trace plane feeds the bulk memory variables (compute_strain_att,
    // compute_element_strain.F90:196: epsilondev_trace = 3 * epsilon_trace_over_3)
    if (with_trace) eps[eps_idx(e, 5, p)] = (realw)3 * templ;
  }
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_elastic)(hipStream_t s, const ElasticArgs& a, bool strain, bool att, int mode, int shape,
                         int att_variant) {
  if (a.nelem <= 0) return;
  const GLLk G = g_gll_host;
  if ((shape == 1 && att && att_variant == 5) || (shape == 2 && att)) {
    const int want = a.pipe_wgs > 0 ? a.pipe_wgs : 2048;
    const int ng = want < a.nelem ? want : a.nelem;
    dim3 grid(ng), block(128);
    if (mode == 1) hipLaunchKernelGGL((k_elastic_pipe<1>), grid, block, 0, s, a, G);
    else if (mode == 2) hipLaunchKernelGGL((k_elastic_pipe<2>), grid, block, 0, s, a, G);
    else hipLaunchKernelGGL((k_elastic_pipe<0>), grid, block, 0, s, a, G);
    SF_LAUNCH_CHECK("k_elastic_pipe");
    return;
  }
  if (shape == 2) {
    // elastic: ten workgroups per CU (90-95 VGPRs, 5 waves per SIMD)
    const int want = a.pipe_wgs > 0 ? a.pipe_wgs : 2560;
    const int ng = want < a.nelem ? want : a.nelem;
    dim3 grid(ng), block(128);
#define SF_KE(S, O) hipLaunchKernelGGL((k_elastic_pipe_el<S, O>), grid, block, 0, s, a, G)
    if (mode == 1) {
      if (strain) SF_KE(true, 1); else SF_KE(false, 1);
    } else if (mode == 2) {
      if (strain) SF_KE(true, 2); else SF_KE(false, 2);
    } else {
      if (strain) SF_KE(true, 0); else SF_KE(false, 0);
    }
#undef SF_KE
    SF_LAUNCH_CHECK("k_elastic_pipe_el");
    return;
  }
  if (shape == 1) {
    dim3 grid(a.nelem), block(128);
#define SF_K(S, T, O, V) hipLaunchKernelGGL((k_elastic_pt<S, T, O, V>), grid, block, 0, s, a, G)
#define SF_KM(O)                                   \
  if (att) {                                       \
    if (att_variant == 1) SF_K(true, true, O, 1);  \
    else if (att_variant == 2) SF_K(true, true, O, 2); \
    else if (att_variant == 3) SF_K(true, true, O, 3); \
    else if (att_variant == 4) SF_K(true, true, O, 4); \
    else SF_K(true, true, O, 0);                   \
  } else if (strain) SF_K(true, false, O, 0);      \
  else SF_K(false, false, O, 0);
    if (mode == 1) {
      SF_KM(1)
    } else if (mode == 2) {
      SF_KM(2)
    } else {
      SF_KM(0)
    }
#undef SF_KM
#undef SF_K
    SF_LAUNCH_CHECK("k_elastic_pt");
    return;
  }
  dim3 grid((a.nelem + 4) / 5), block(128);
#define SF_K(S, T, O) hipLaunchKernelGGL((k_elastic<S, T, O>), grid, block, 0, s, a, G)
#define SF_KM(O)                      \
  if (att) SF_K(true, true, O);       \
  else if (strain) SF_K(true, false, O); \
  else SF_K(false, false, O);
  if (mode == 1) {
    SF_KM(1)
  } else if (mode == 2) {
    SF_KM(2)
  } else {
    SF_KM(0)
  }
#undef SF_KM
#undef SF_K
  SF_LAUNCH_CHECK("k_elastic");
}

void SFX(launch_element_strain)(hipStream_t s, const ElasticArgs& a, realw* eps, realw* eps_tr3, bool with_trace) {
  if (a.nelem <= 0) return;
  dim3 grid((a.nelem + 4) / 5), block(128);
  hipLaunchKernelGGL(k_element_strain, grid, block, 0, s, a, g_gll_host, eps, eps_tr3, with_trace ? 1 : 0);
  SF_LAUNCH_CHECK("k_element_strain");
}

}  // namespace sf
