// k_elastic.hip -- isotropic elastic / viscoelastic (3 SLS) element stiffness for NGLL=5.
//
// Restates compute_forces_viscoelastic.F90:33-962 (with the memory update of
// compute_element_att_memory.f90:28-156) as one gfx950 kernel, k_elastic_pt:
//   * a 128-thread workgroup (two wave64s) processes one element; lane p owns the GLL point
//     p = i + 5 j + 25 k (125 of the 128 lanes busy) and all three contractions read the element
//     tile staged in LDS;
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
// gathered.
// AV selects how the attenuation state is scheduled (the "att_variant" tuning switch; same
// results):
// 0 every per-point array is loaded before the first barrier; 1 only the memory variables are,
// the stored strain and SLS factors are loaded where the memory update needs them (fewer live
// registers); 2 as 1 with the register budget capped for 5 waves per SIMD; 3 as 0 with that cap;
// 4 loads the memory variables and stored strain straight into LDS (global_load_lds, no
// registers held while they are in flight; each lane reads back its own point), which frees
// registers for 5 waves per SIMD; the stress products then reuse those LDS planes.
typedef __attribute__((address_space(3))) void* sf_lds_ptr;
typedef __attribute__((address_space(1))) void* sf_global_ptr;

template <bool STRAIN, bool ATT, int MODE, int AV>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(AV >= 2 ? 5 : 1)))
k_elastic_pt(ElasticArgs A, GLLk G) {
  constexpr bool LATE = ATT && (AV == 1 || AV == 2);
  constexpr bool DL = ATT && AV == 4;
  __shared__ realw sU[6][128];            // u (0..2) and dt*v (3..5)
  __shared__ realw sX[DL ? 24 : 9][128];  // stress products (planes 0..8); with DL first the state
  realw(*const sT)[128] = sX;             // t1 (0..2), t2 (3..5), t3 (6..8)
  __shared__ realw sG[SF_NGLL_TABLE];     // hprime, hprimewgll, wgllwgll (+ folded, elastic_point.h)
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
  const realw* hpg = gll_hp(sG, irr != 0);
  const Grad9 tg = tile_gradient<1>(sU[0], sU[1], sU[2], hpg, i, j, k);
  const Du9 du = map_gradient(tg, irr != 0, J, A.xix_regular);

  realw e_xx = 0, e_yy = 0, e_xy = 0, e_xz = 0, e_yz = 0, e_tr = 0;
  if (STRAIN) {
    Du9 st = du;
    if (ATT) st = map_gradient(tile_gradient_continue<1>(tg, sU[3], sU[4], sU[5], hpg, i, j, k), irr != 0, J,
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
  tile_divergence<1>(&sT[0][0], 128, gll_hpw(sG, irr != 0), sG + 50, i, j, k, f);
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
    // (plain stores of the 27 element-interior points, which belong to this element alone, measured
    // 16 % slower on the 96x96x48 viscoelastic workload than atomics for all points)
    realw* a = A.accel + 3 * (int64_t)g;
    atomicAdd(a + 0, -f[0]);
    atomicAdd(a + 1, -f[1]);
    atomicAdd(a + 2, -f[2]);
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

void SFX(launch_elastic)(hipStream_t s, const ElasticArgs& a, bool strain, bool att, int mode, int att_variant) {
  if (a.nelem <= 0) return;
  GLLk G = g_gll_host;
  if (SF_FOLD) fold_gll(G, a.xix_regular, a.jacobian_regular);
  dim3 grid(a.nelem), block(128);
#define SF_K(S, T, O, V) hipLaunchKernelGGL((k_elastic_pt<S, T, O, V>), grid, block, 0, s, a, G)
#define SF_KM(O)                                       \
  if (att) {                                           \
    if (att_variant == 1) SF_K(true, true, O, 1);      \
    else if (att_variant == 2) SF_K(true, true, O, 2); \
    else if (att_variant == 3) SF_K(true, true, O, 3); \
    else if (att_variant == 4) SF_K(true, true, O, 4); \
    else SF_K(true, true, O, 0);                       \
  } else if (strain) SF_K(true, false, O, 0);          \
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
}

void SFX(launch_element_strain)(hipStream_t s, const ElasticArgs& a, realw* eps, realw* eps_tr3, bool with_trace) {
  if (a.nelem <= 0) return;
  dim3 grid((a.nelem + 4) / 5), block(128);
  hipLaunchKernelGGL(k_element_strain, grid, block, 0, s, a, g_gll_host, eps, eps_tr3, with_trace ? 1 : 0);
  SF_LAUNCH_CHECK("k_element_strain");
}

}  // namespace sf
