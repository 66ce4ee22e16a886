// k_pml.hip -- C-PML elements of the elastic stiffness for NGLL=5.
//
// Restates compute_forces_viscoelastic_PML (compute_forces_viscoelastic.F90:969-1748) with its
// memory-variable update pml_compute_memory_variables_elastic (pml_compute_memory_variables.f90:
// 29-415) and the damping term pml_compute_accel_contribution_elastic
// (pml_compute_accel_contribution.f90:29-146), as one point-per-lane kernel like k_elastic_pt:
//   * a 128-thread workgroup processes one C-PML element, lane p the GLL point p = i + 5 j + 25 k;
//   * three element tiles sit in LDS: u, PML_displ_old and PML_displ_new. PML_displ_old is read per
//     global point (it depends only on the point's u, v, a before the predictor); PML_displ_new is
//     recomputed from u and v (the predictor leaves a = 0), so neither is stored per local point;
//   * the 48 memory variables and 42 coefficients of a point are 125-point planes: each access of a
//     wavefront is one coalesced run. Per element, coefficient planes holding one value are read at
//     one address, memory variables whose multiplier is zero over the element (most of them in
//     face and edge regions) are not kept at all, and the remaining planes are packed;
//   * the scatter is selected by MODE: 0 float atomics, 1 a local store for the ordered gather
//     (one launch in both cases: the C-PML elements alone are too few to fill the chip colour by
//     colour).
// Every sum keeps the reference's operand order; compiled twice (SF_EXACT = 1: no FMA contraction,
// the CPU reference's rounding; SF_EXACT = 0: FMA, with the regular-element mapping folded into
// the GLL tables as in k_elastic_pt).
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
namespace SF_VNS {

__device__ __forceinline__ int pml_xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// r <- coef0 r + new coef1 + old coef2: the second-order recursive convolution of
// pml_compute_memory_variables.f90:152-296 / pml_compute_accel_contribution.f90:88-121
__device__ __forceinline__ realw pml_conv(realw r, realw nw, realw od, realw c0, realw c1, realw c2) {
  return c0 * r + nw * c1 + od * c2;
}

// coefficient planes (before packing): alpha_x 0..2, alpha_y 3..5, alpha_z 6..8 (coef0, coef1,
// coef2), beta likewise at 9..17, A_1..A_5 at 18..22, A_6..A_23 at 23..40, wgll_cube * rho at 41.
// state planes: three-term memory variables (a * 3 + s) for a = dux_dxl_x, duy_dxl_y, duz_dxl_z,
// dux_dyl_x, duy_dyl_y, duz_dyl_z, dux_dzl_x, duy_dzl_y, duz_dzl_z at 0..26; one-term ones at 27..38
// (duz_dzl_y, duz_dyl_y, duy_dzl_z, duy_dyl_z | duz_dzl_x, duz_dxl_x, dux_dzl_z, dux_dxl_z |
// duy_dyl_x, duy_dxl_x, dux_dyl_y, dux_dxl_y); displacement memory variables (c * 3 + s) at 39..47.
template <int MODE>
__global__ void __launch_bounds__(128) k_elastic_pml(PmlArgs A, GLLk G) {
  __shared__ realw sF[9][128];  // u (0..2), PML_displ_old (3..5), PML_displ_new (6..8)
  __shared__ realw sT[9][128];  // stress products t1 (0..2), t2 (3..5), t3 (6..8)
  __shared__ realw sG[SF_NGLL_TABLE];
  const int p = threadIdx.x;
  const bool active = p < 125;
  stage_gll(sG, G, p);
  const int ie = pml_xcd_remap(blockIdx.x, gridDim.x);
  const int e = A.elems[ie];
  const int64_t ic = A.cpml[ie];
  const int q = active ? p : 124;
  const int k = q / 25, j = (q / 5) % 5, i = q % 5;
  const int64_t ep = (int64_t)e * 125 + q;
  const int g = A.ibool[ep];
  const int irr = A.irregular[e];
  JacPt J = {};
  realw jacl = A.jacobian_regular;
  if (irr) {
    J = load_jac(A.jac, A.jac_plane, (int64_t)(irr - 1) * 125 + q);
    jacl = J.jac;
  }
  realw u[3], v[3], od[3], nw[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    u[c] = A.displ[3 * (int64_t)g + c];
    v[c] = A.veloc[3 * (int64_t)g + c];
    od[c] = A.pml_old[3 * (int64_t)g + c];
    // update_displ_elastic_PML after the predictor, which has set accel to zero
    nw[c] = (u[c] + A.t1 * v[c]) + A.t2 * (realw)0;
    sF[c][p] = u[c];
    sF[3 + c][p] = od[c];
    sF[6 + c][p] = nw[c];
  }
  __syncthreads();

  const realw* hpg = gll_hp(sG, irr != 0);
  const Du9 C = map_gradient(tile_gradient<1>(sF[0], sF[1], sF[2], hpg, i, j, k), irr != 0, J, A.xix_regular);
  const Du9 O = map_gradient(tile_gradient<1>(sF[3], sF[4], sF[5], hpg, i, j, k), irr != 0, J, A.xix_regular);
  const Du9 N = map_gradient(tile_gradient<1>(sF[6], sF[7], sF[8], hpg, i, j, k), irr != 0, J, A.xix_regular);

  // element-uniform masks and packed planes: the coefficient planes of an element sit in plane
  // order at its offset, a uniform one as its single value (one address per wavefront), the others
  // as 125 values; the live state planes likewise, 125 values each. A plane's position is counted
  // from the mask bits below it; dead state planes enter as zero (their multipliers are zero)
  const uint64_t um = A.masks[2 * ic], lv = A.masks[2 * ic + 1];
  const realw* cf = A.coef + A.off[2 * ic];
  realw* st = A.state + A.off[2 * ic + 1];
#define BELOW(m) ((1ull << (m)) - 1ull)
#define CF(m) cf[__builtin_popcountll(~um & BELOW(m)) * 125 + __builtin_popcountll(um & BELOW(m)) + \
                 (((um >> (m)) & 1) ? 0 : q)]
#define LIVE(pl) ((lv >> (pl)) & 1)
#define ST(pl) st[__builtin_popcountll(lv & BELOW(pl)) * 125 + q]

  // three-term memory variables: derivative direction d, component c; coefficient sets
  // (alpha_z, alpha_y, beta_x) for d/dx, (alpha_x, alpha_z, beta_y) for d/dy, (alpha_x, alpha_y, beta_z) for d/dz
  const realw nv[9] = {N.xx, N.yx, N.zx, N.xy, N.yy, N.zy, N.xz, N.yz, N.zz};
  const realw ov[9] = {O.xx, O.yx, O.zx, O.xy, O.yy, O.zy, O.xz, O.yz, O.zz};
  constexpr int cset[3][3] = {{6, 3, 9}, {0, 6, 12}, {0, 3, 15}};
  realw r3[9][3];
#pragma unroll
  for (int a = 0; a < 9; a++) {
    const int d = a / 3;
#pragma unroll
    for (int s = 0; s < 3; s++) {
      const int b = cset[d][s];
      const int pl = a * 3 + s;
      realw r = 0;
      if (LIVE(pl)) {
        r = pml_conv(ST(pl), nv[a], ov[a], CF(b), CF(b + 1), CF(b + 2));
        if (active) ST(pl) = r;
      }
      r3[a][s] = r;
    }
  }
  // one-term memory variables (alpha_x for 0..3, alpha_y for 4..7, alpha_z for 8..11)
  const realw n1[12] = {N.zz, N.zy, N.yz, N.yy, N.zz, N.zx, N.xz, N.xx, N.yy, N.yx, N.xy, N.xx};
  const realw o1[12] = {O.zz, O.zy, O.yz, O.yy, O.zz, O.zx, O.xz, O.xx, O.yy, O.yx, O.xy, O.xx};
  realw r1[12];
#pragma unroll
  for (int b = 0; b < 12; b++) {
    const int cb = 3 * (b / 4);
    realw r = 0;
    if (LIVE(27 + b)) {
      r = pml_conv(ST(27 + b), n1[b], o1[b], CF(cb), CF(cb + 1), CF(cb + 2));
      if (active) ST(27 + b) = r;
    }
    r1[b] = r;
  }
  // strain of the PML stress (A_6..A_23 = plane 23 + 0..17)
#define AS(m) CF(23 + (m))
  const realw duxdxl_x = AS(0) * C.xx + AS(1) * r3[0][0] + AS(2) * r3[0][1] + AS(3) * r3[0][2];
  const realw duxdyl_x = AS(4) * C.xy + AS(5) * r3[3][0] + AS(6) * r3[3][1] + AS(7) * r3[3][2];
  const realw duxdzl_x = AS(8) * C.xz + AS(9) * r3[6][0] + AS(10) * r3[6][1] + AS(11) * r3[6][2];
  const realw duzdzl_x = AS(14) * C.zz + AS(15) * r1[4];
  const realw duzdxl_x = AS(14) * C.zx + AS(15) * r1[5];
  const realw duydyl_x = AS(16) * C.yy + AS(17) * r1[8];
  const realw duydxl_x = AS(16) * C.yx + AS(17) * r1[9];
  const realw duydxl_y = AS(0) * C.yx + AS(1) * r3[1][0] + AS(2) * r3[1][1] + AS(3) * r3[1][2];
  const realw duydyl_y = AS(4) * C.yy + AS(5) * r3[4][0] + AS(6) * r3[4][1] + AS(7) * r3[4][2];
  const realw duydzl_y = AS(8) * C.yz + AS(9) * r3[7][0] + AS(10) * r3[7][1] + AS(11) * r3[7][2];
  const realw duzdzl_y = AS(12) * C.zz + AS(13) * r1[0];
  const realw duzdyl_y = AS(12) * C.zy + AS(13) * r1[1];
  const realw duxdyl_y = AS(16) * C.xy + AS(17) * r1[10];
  const realw duxdxl_y = AS(16) * C.xx + AS(17) * r1[11];
  const realw duzdxl_z = AS(0) * C.zx + AS(1) * r3[2][0] + AS(2) * r3[2][1] + AS(3) * r3[2][2];
  const realw duzdyl_z = AS(4) * C.zy + AS(5) * r3[5][0] + AS(6) * r3[5][1] + AS(7) * r3[5][2];
  const realw duzdzl_z = AS(8) * C.zz + AS(9) * r3[8][0] + AS(10) * r3[8][1] + AS(11) * r3[8][2];
  const realw duydzl_z = AS(12) * C.yz + AS(13) * r1[2];
  const realw duydyl_z = AS(12) * C.yy + AS(13) * r1[3];
  const realw duxdzl_z = AS(14) * C.xz + AS(15) * r1[6];
  const realw duxdxl_z = AS(14) * C.xx + AS(15) * r1[7];
#undef AS
  // non-symmetric PML stress (pml_compute_memory_variables.f90:370-385)
  const float2 km = A.kappamu[ep];
  const realw kappal = km.x, mul = km.y;
  const realw FOUR_THIRDS = (realw)4 / (realw)3;
  const realw l2m = kappal + FOUR_THIRDS * mul;
  const realw lam = l2m - (realw)2 * mul;
  const realw sxx = l2m * duxdxl_x + lam * duydyl_x + lam * duzdzl_x;
  const realw syx = mul * duxdyl_x + mul * duydxl_x;
  const realw szx = mul * duzdxl_x + mul * duxdzl_x;
  const realw sxy = mul * duxdyl_y + mul * duydxl_y;
  const realw syy = lam * duxdxl_y + l2m * duydyl_y + lam * duzdzl_y;
  const realw szy = mul * duzdyl_y + mul * duydzl_y;
  const realw sxz = mul * duzdxl_z + mul * duxdzl_z;
  const realw syz = mul * duzdyl_z + mul * duydzl_z;
  const realw szz = lam * duxdxl_z + lam * duydyl_z + l2m * duzdzl_z;
  {
    realw t[9];
    if (irr) {
      t[0] = jacl * (sxx * J.xix + syx * J.xiy + szx * J.xiz);
      t[1] = jacl * (sxy * J.xix + syy * J.xiy + szy * J.xiz);
      t[2] = jacl * (sxz * J.xix + syz * J.xiy + szz * J.xiz);
      t[3] = jacl * (sxx * J.etax + syx * J.etay + szx * J.etaz);
      t[4] = jacl * (sxy * J.etax + syy * J.etay + szy * J.etaz);
      t[5] = jacl * (sxz * J.etax + syz * J.etay + szz * J.etaz);
      t[6] = jacl * (sxx * J.gax + syx * J.gay + szx * J.gaz);
      t[7] = jacl * (sxy * J.gax + syy * J.gay + szy * J.gaz);
      t[8] = jacl * (sxz * J.gax + syz * J.gay + szz * J.gaz);
    } else if (SF_FOLD) {
      t[0] = sxx; t[1] = sxy; t[2] = sxz;
      t[3] = syx; t[4] = syy; t[5] = syz;
      t[6] = szx; t[7] = szy; t[8] = szz;
    } else {
      const realw xr = A.xix_regular;
      t[0] = jacl * sxx * xr; t[1] = jacl * sxy * xr; t[2] = jacl * sxz * xr;
      t[3] = jacl * syx * xr; t[4] = jacl * syy * xr; t[5] = jacl * syz * xr;
      t[6] = jacl * szx * xr; t[7] = jacl * szy * xr; t[8] = jacl * szz * xr;
    }
#pragma unroll
    for (int m = 0; m < 9; m++) sT[m][p] = t[m];
  }
  // damping term of the acceleration (pml_compute_accel_contribution_elastic)
  realw cpml[3];
  {
    const realw w = CF(41) * jacl;  // wgll_cube * rho * jacobian
    const realw A1 = CF(18), A2 = CF(19), A3 = CF(20), A4 = CF(21), A5 = CF(22);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      realw rr[3];
#pragma unroll
      for (int s = 0; s < 3; s++) {
        const int b = 3 * s;
        const int pl = 39 + 3 * c + s;
        realw r = 0;
        if (LIVE(pl)) {
          r = pml_conv(ST(pl), nw[c], od[c], CF(b), CF(b + 1), CF(b + 2));
          if (active) ST(pl) = r;
        }
        rr[s] = r;
      }
      cpml[c] = w * (A1 * v[c] + A2 * u[c] + A3 * rr[0] + A4 * rr[1] + A5 * rr[2]);
    }
  }
#undef CF
#undef LIVE
#undef ST
#undef BELOW
  lds_barrier();
  if (!active) return;
  realw f[3];
  tile_divergence<1>(&sT[0][0], 128, gll_hpw(sG, irr != 0), sG + 50, i, j, k, f);
  const realw tot0 = f[0] + cpml[0], tot1 = f[1] + cpml[1], tot2 = f[2] + cpml[2];
  if (MODE == 1) {
    A.local_out[ep] = tot0;
    A.local_out[A.nloc + ep] = tot1;
    A.local_out[2 * A.nloc + ep] = tot2;
  } else {
    realw* a = A.accel + 3 * (int64_t)g;
    atomicAdd(a + 0, -tot0);
    atomicAdd(a + 1, -tot1);
    atomicAdd(a + 2, -tot2);
  }
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_elastic_pml)(hipStream_t s, const PmlArgs& a, int mode) {
  if (a.nelem <= 0) return;
  GLLk G = g_gll_host;
  if (SF_FOLD) fold_gll(G, a.xix_regular, a.jacobian_regular);
  dim3 grid(a.nelem), block(128);
  if (mode == 1) hipLaunchKernelGGL(k_elastic_pml<1>, grid, block, 0, s, a, G);
  else hipLaunchKernelGGL(k_elastic_pml<0>, grid, block, 0, s, a, G);
  SF_LAUNCH_CHECK("k_elastic_pml");
}

}  // namespace sf
