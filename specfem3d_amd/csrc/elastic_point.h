// elastic_point.h -- per-GLL-point arithmetic of the isotropic elastic stiffness, shared by the
// single-shot point-per-lane kernel (k_elastic.hip) and the batched many-shot kernel
// (k_batch.hip). Every expression keeps the reference's operand order
// (compute_forces_viscoelastic.F90:331-335 contractions, :420-461 mapping, :616-630 stress,
// :661-710 test-function products, :722-726 divergence, :788-806 weights), so the exact build
// (-ffp-contract=off) reproduces the CPU solver bit for bit whichever kernel runs it.
//
// Element tiles live in LDS with the value of point q = i + 5 j + 25 k at base[q * PS]: PS = 1 for
// one field per workgroup, PS = shots per workgroup for shot-interleaved tiles. The GLL tables
// (hp[i*5+l] = hprime, hpw[i*5+l] = hprimewgll, wxy[a*5+b] = wgllwgll) are read through pointers:
// the kernels stage them in LDS so a lane reads each coefficient when it uses it instead of
// holding its 33 coefficients in registers.
#pragma once

#include "kernels.h"

#ifndef SF_EXACT
#define SF_EXACT 1
#endif
// FMA builds fold, on regular elements, xix_regular into hprime and jacobian_regular *
// xix_regular into hprimewgll (the tables at sG + 75 and sG + 100): the nine-term mapping of the
// gradient and the eighteen multiplies of the test-function products become copies. The exact
// builds keep the reference's operation order.
#define SF_FOLD (!SF_EXACT)

namespace sf {

// Workgroup barrier that orders LDS accesses only: this wave's LDS operations complete
// (lgkmcnt(0)), then s_barrier; the "memory" clobber keeps the compiler from moving memory
// accesses across it. Global loads, stores and atomics issued earlier, and LDS-direct loads
// (global_load_lds, which complete through vmcnt and are read back only by the lane that issued
// them), stay in flight across it; __syncthreads would wait for all of them (vmcnt(0)).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct Grad9 {  // d/dxi, d/deta, d/dgamma of the three components
  realw x1, y1, z1, x2, y2, z2, x3, y3, z3;
};

struct JacPt {  // inverse-Jacobian terms and determinant at one point (irregular elements)
  realw xix, xiy, xiz, etax, etay, etaz, gax, gay, gaz, jac;
};

// tempx1 = sum_l hprime(i,l) u(l,j,k), tempx2 = sum_l hprime(j,l) u(i,l,k),
// tempx3 = sum_l hprime(k,l) u(i,j,l), summed left to right.
template <int PS>
__device__ __forceinline__ Grad9 tile_gradient(const realw* sx, const realw* sy, const realw* sz, const realw* hp,
                                               int i, int j, int k) {
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
  Grad9 t;
  const realw h1 = hp[i * 5], h2 = hp[j * 5], h3 = hp[k * 5];
  t.x1 = h1 * sx[b1 * PS]; t.y1 = h1 * sy[b1 * PS]; t.z1 = h1 * sz[b1 * PS];
  t.x2 = h2 * sx[b2 * PS]; t.y2 = h2 * sy[b2 * PS]; t.z2 = h2 * sz[b2 * PS];
  t.x3 = h3 * sx[b3 * PS]; t.y3 = h3 * sy[b3 * PS]; t.z3 = h3 * sz[b3 * PS];
#pragma unroll
  for (int l = 1; l < 5; l++) {
    const realw a1 = hp[i * 5 + l], a2 = hp[j * 5 + l], a3 = hp[k * 5 + l];
    const int q1 = (b1 + l) * PS, q2 = (b2 + 5 * l) * PS, q3 = (b3 + 25 * l) * PS;
    t.x1 = t.x1 + a1 * sx[q1]; t.y1 = t.y1 + a1 * sy[q1]; t.z1 = t.z1 + a1 * sz[q1];
    t.x2 = t.x2 + a2 * sx[q2]; t.y2 = t.y2 + a2 * sy[q2]; t.z2 = t.z2 + a2 * sz[q2];
    t.x3 = t.x3 + a3 * sx[q3]; t.y3 = t.y3 + a3 * sy[q3]; t.z3 = t.z3 + a3 * sz[q3];
  }
  return t;
}

// Strain at u + dt v (attenuation, compute_forces_viscoelastic.F90:396-417): each contraction of u
// is continued with the five terms of the dt*v tile.
template <int PS>
__device__ __forceinline__ Grad9 tile_gradient_continue(Grad9 t, const realw* wx, const realw* wy, const realw* wz,
                                                        const realw* hp, int i, int j, int k) {
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const realw h1 = hp[i * 5 + l], h2 = hp[j * 5 + l], h3 = hp[k * 5 + l];
    const int q1 = (b1 + l) * PS, q2 = (b2 + 5 * l) * PS, q3 = (b3 + 25 * l) * PS;
    t.x1 = t.x1 + wx[q1] * h1; t.y1 = t.y1 + wy[q1] * h1; t.z1 = t.z1 + wz[q1] * h1;
    t.x2 = t.x2 + wx[q2] * h2; t.y2 = t.y2 + wy[q2] * h2; t.z2 = t.z2 + wz[q2] * h2;
    t.x3 = t.x3 + wx[q3] * h3; t.y3 = t.y3 + wy[q3] * h3; t.z3 = t.z3 + wz[q3] * h3;
  }
  return t;
}

// The nine displacement derivatives in (x, y, z): d(ux)/dx .. d(uz)/dz.
struct Du9 {
  realw xx, xy, xz, yx, yy, yz, zx, zy, zz;
};

__device__ __forceinline__ Du9 map_gradient(const Grad9& t, bool irr, const JacPt& J, realw xr) {
  Du9 d;
  if (irr) {
    d.xx = J.xix * t.x1 + J.etax * t.x2 + J.gax * t.x3;
    d.xy = J.xiy * t.x1 + J.etay * t.x2 + J.gay * t.x3;
    d.xz = J.xiz * t.x1 + J.etaz * t.x2 + J.gaz * t.x3;
    d.yx = J.xix * t.y1 + J.etax * t.y2 + J.gax * t.y3;
    d.yy = J.xiy * t.y1 + J.etay * t.y2 + J.gay * t.y3;
    d.yz = J.xiz * t.y1 + J.etaz * t.y2 + J.gaz * t.y3;
    d.zx = J.xix * t.z1 + J.etax * t.z2 + J.gax * t.z3;
    d.zy = J.xiy * t.z1 + J.etay * t.z2 + J.gay * t.z3;
    d.zz = J.xiz * t.z1 + J.etaz * t.z2 + J.gaz * t.z3;
  } else if (SF_FOLD) {
    d.xx = t.x1; d.xy = t.x2; d.xz = t.x3;
    d.yx = t.y1; d.yy = t.y2; d.yz = t.y3;
    d.zx = t.z1; d.zy = t.z2; d.zz = t.z3;
  } else {
    d.xx = xr * t.x1; d.xy = xr * t.x2; d.xz = xr * t.x3;
    d.yx = xr * t.y1; d.yy = xr * t.y2; d.yz = xr * t.y3;
    d.zx = xr * t.z1; d.zy = xr * t.z2; d.zz = xr * t.z3;
  }
  return d;
}

struct Sigma6 {
  realw xx, yy, zz, xy, xz, yz;
};

// isotropic stress, lambda + 2 mu = kappa + 4/3 mu, lambda = (lambda + 2 mu) - 2 mu
__device__ __forceinline__ Sigma6 iso_stress(const Du9& d, realw kappal, realw mul) {
  const realw FOUR_THIRDS = (realw)4 / (realw)3;
  const realw lp2m = kappal + FOUR_THIRDS * mul;
  const realw lambdal = lp2m - (realw)2 * mul;
  Sigma6 s;
  s.xx = lp2m * d.xx + lambdal * (d.yy + d.zz);
  s.yy = lp2m * d.yy + lambdal * (d.xx + d.zz);
  s.zz = lp2m * d.zz + lambdal * (d.xx + d.yy);
  s.xy = mul * (d.xy + d.yx);
  s.xz = mul * (d.zx + d.xz);
  s.yz = mul * (d.zy + d.yz);
  return s;
}

// jacobian * sigma . (dxi/dx), (deta/dx), (dgamma/dx): t[0..2] xi, t[3..5] eta, t[6..8] gamma
__device__ __forceinline__ void stress_products(const Sigma6& s, bool irr, const JacPt& J, realw jacl, realw xr,
                                                realw t[9]) {
  const realw syx = s.xy, szx = s.xz, szy = s.yz;
  if (irr) {
    t[0] = jacl * (s.xx * J.xix + syx * J.xiy + szx * J.xiz);
    t[1] = jacl * (s.xy * J.xix + s.yy * J.xiy + szy * J.xiz);
    t[2] = jacl * (s.xz * J.xix + s.yz * J.xiy + s.zz * J.xiz);
    t[3] = jacl * (s.xx * J.etax + syx * J.etay + szx * J.etaz);
    t[4] = jacl * (s.xy * J.etax + s.yy * J.etay + szy * J.etaz);
    t[5] = jacl * (s.xz * J.etax + s.yz * J.etay + s.zz * J.etaz);
    t[6] = jacl * (s.xx * J.gax + syx * J.gay + szx * J.gaz);
    t[7] = jacl * (s.xy * J.gax + s.yy * J.gay + szy * J.gaz);
    t[8] = jacl * (s.xz * J.gax + s.yz * J.gay + s.zz * J.gaz);
  } else if (SF_FOLD) {
    t[0] = s.xx; t[1] = s.xy; t[2] = s.xz;
    t[3] = syx;  t[4] = s.yy; t[5] = s.yz;
    t[6] = szx;  t[7] = szy;  t[8] = s.zz;
  } else {
    t[0] = jacl * s.xx * xr;
    t[1] = jacl * s.xy * xr;
    t[2] = jacl * s.xz * xr;
    t[3] = jacl * syx * xr;
    t[4] = jacl * s.yy * xr;
    t[5] = jacl * s.yz * xr;
    t[6] = jacl * szx * xr;
    t[7] = jacl * szy * xr;
    t[8] = jacl * s.zz * xr;
  }
}

// f_c = wgllwgll_yz(j,k) sum_l hprimewgll(l,i) t1_c(l,j,k) + wgllwgll_xz(i,k) sum_l hprimewgll(l,j) t2_c(i,l,k)
//     + wgllwgll_xy(i,j) sum_l hprimewgll(l,k) t3_c(i,j,l); plane m of the tile at sT + m * plane.
template <int PS>
__device__ __forceinline__ void tile_divergence(const realw* sT, int plane, const realw* hpw, const realw* wxy, int i,
                                                int j, int k, realw f[3]) {
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
#pragma unroll
  for (int cc = 0; cc < 3; cc++) {
    const realw *s1 = sT + cc * plane, *s2 = sT + (3 + cc) * plane, *s3 = sT + (6 + cc) * plane;
    realw n1 = hpw[i] * s1[b1 * PS];
    realw n2 = hpw[j] * s2[b2 * PS];
    realw n3 = hpw[k] * s3[b3 * PS];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      n1 = n1 + hpw[l * 5 + i] * s1[(b1 + l) * PS];
      n2 = n2 + hpw[l * 5 + j] * s2[(b2 + 5 * l) * PS];
      n3 = n3 + hpw[l * 5 + k] * s3[(b3 + 25 * l) * PS];
    }
    f[cc] = wxy[j * 5 + k] * n1 + wxy[i * 5 + k] * n2 + wxy[i * 5 + j] * n3;
  }
}

// one component cc of tile_divergence (same expressions)
template <int PS>
__device__ __forceinline__ realw tile_divergence_c(const realw* sT, int plane, const realw* hpw, const realw* wxy,
                                                   int i, int j, int k, int cc) {
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
  const realw *s1 = sT + cc * plane, *s2 = sT + (3 + cc) * plane, *s3 = sT + (6 + cc) * plane;
  realw n1 = hpw[i] * s1[b1 * PS];
  realw n2 = hpw[j] * s2[b2 * PS];
  realw n3 = hpw[k] * s3[b3 * PS];
#pragma unroll
  for (int l = 1; l < 5; l++) {
    n1 = n1 + hpw[l * 5 + i] * s1[(b1 + l) * PS];
    n2 = n2 + hpw[l * 5 + j] * s2[(b2 + 5 * l) * PS];
    n3 = n3 + hpw[l * 5 + k] * s3[(b3 + 25 * l) * PS];
  }
  return wxy[j * 5 + k] * n1 + wxy[i * 5 + k] * n2 + wxy[i * 5 + j] * n3;
}

// stage the GLL coefficients of G in LDS (lanes 0..124 of the workgroup; needs a barrier before
// use): hp, hpw, wxy, and the folded hpr, hpwr at sG + 75, sG + 100
constexpr int SF_NGLL_TABLE = 125;
__device__ __forceinline__ void stage_gll(realw* sG, const GLLk& G, int lane) {
  if (lane < 25) sG[lane] = G.hp[lane];
  else if (lane < 50) sG[lane] = G.hpw[lane - 25];
  else if (lane < 75) sG[lane] = G.wxy[lane - 50];
  else if (lane < 100) sG[lane] = G.hpr[lane - 75];
  else if (lane < 125) sG[lane] = G.hpwr[lane - 100];
}
// the hprime / hprimewgll table an element uses (folded ones on regular elements in FMA builds)
__device__ __forceinline__ const realw* gll_hp(const realw* sG, bool irr) { return (SF_FOLD && !irr) ? sG + 75 : sG; }
__device__ __forceinline__ const realw* gll_hpw(const realw* sG, bool irr) {
  return (SF_FOLD && !irr) ? sG + 100 : sG + 25;
}

__device__ __forceinline__ JacPt load_jac(const realw* J, int64_t P, int64_t q) {
  JacPt r;
  r.xix = J[q]; r.xiy = J[P + q]; r.xiz = J[2 * P + q];
  r.etax = J[3 * P + q]; r.etay = J[4 * P + q]; r.etaz = J[5 * P + q];
  r.gax = J[6 * P + q]; r.gay = J[7 * P + q]; r.gaz = J[8 * P + q];
  r.jac = J[9 * P + q];
  return r;
}

}  // namespace sf
