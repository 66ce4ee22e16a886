// k_batch.hip -- element stiffness of the batched many-shot engine (elastic, isotropic).
//
// Many independent shots share one mesh. Their wavefields are interleaved shot-fastest: the value
// (global point g, component c, shot s) lives at ((int64_t)g * 3 + c) * S + s, S = padded shot
// count. A workgroup of 256 lanes takes one element and a group of SC shots: lane = (point slot,
// shot) with the shot fastest, so every gather/scatter of a point row is SC consecutive floats, and
// the element's ibool, moduli and Jacobian terms are loaded once for the SC shots (the lanes of one
// point read the same address). Workgroups of one element's shot groups are adjacent in the grid,
// so the groups share the L2 lines of each row. The per-point arithmetic is elastic_point.h, the
// same code as the single-shot kernel, so a batched shot reproduces its single-shot run.
//
// Shots are stored in decreasing order of record length; a shot is live at step `it` while
// nstep[s] >= it, and a group whose first shot has finished exits at once (uneven record
// lengths cost nothing after a shot ends). `it` is read from device memory so a captured graph
// of the step can be replayed.
//
// Compiled twice like k_elastic.hip: SF_EXACT=1 without FMA contraction, SF_EXACT=0 with it.
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

__device__ __forceinline__ int xcd_remap_b(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

template <int SC, int MODE>
__global__ void __launch_bounds__(256) k_elastic_batch(BatchArgs B, GLLk G) {
  constexpr int NP = 256 / SC;                    // point slots per pass
  constexpr int NPASS = (125 + NP - 1) / NP;
  constexpr int PL = 125 * SC;                    // one plane of the tile
  __shared__ realw sU[3 * PL];                    // displacement tile [c][p][s]
  __shared__ realw sT[9 * PL];                    // stress products   [m][p][s]
  __shared__ realw sG[75];                        // hprime, hprimewgll, wgllwgll
  stage_gll(sG, G, threadIdx.x);
  const int it = __hip_atomic_load(B.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int b = xcd_remap_b(blockIdx.x, gridDim.x);
  const int ie = b / B.ngroups;
  const int s0 = (b - ie * B.ngroups) * SC;
  if (B.nstep[s0] < it) return;                   // every shot of this group has finished
  const int sl = threadIdx.x % SC, pl = threadIdx.x / SC;
  const int s = s0 + sl;
  const bool live = B.nstep[s] >= it;
  const int e = B.elems[ie];
  const int irr = B.irregular[e];
  const int64_t S = B.S;

  int gq[NPASS];
#pragma unroll
  for (int ps = 0; ps < NPASS; ps++) {
    const int p = pl + ps * NP;
    gq[ps] = 0;
    if (p < 125) {
      const int g = B.ibool[(int64_t)e * 125 + p];
      gq[ps] = g;
      const int64_t row = (int64_t)g * 3 * S + s;
      sU[(0 * 125 + p) * SC + sl] = live ? B.u[row] : (realw)0;
      sU[(1 * 125 + p) * SC + sl] = live ? B.u[row + S] : (realw)0;
      sU[(2 * 125 + p) * SC + sl] = live ? B.u[row + 2 * S] : (realw)0;
    }
  }
  __syncthreads();

#pragma unroll
  for (int ps = 0; ps < NPASS; ps++) {
    const int p = pl + ps * NP;
    if (p < 125) {
      const int k = p / 25, j = (p / 5) % 5, i = p % 5;
      const Grad9 tg = tile_gradient<SC>(sU + sl, sU + PL + sl, sU + 2 * PL + sl, sG, i, j, k);
      JacPt J = {};
      realw jacl = B.jacobian_regular;
      if (irr) {
        J = load_jac(B.jac, B.jac_plane, (int64_t)(irr - 1) * 125 + p);
        jacl = J.jac;
      }
      const Du9 du = map_gradient(tg, irr != 0, J, B.xix_regular);
      const float2 km = B.kappamu[(int64_t)e * 125 + p];
      const Sigma6 sg = iso_stress(du, km.x, km.y);
      realw t[9];
      stress_products(sg, irr != 0, J, jacl, B.xix_regular, t);
#pragma unroll
      for (int m = 0; m < 9; m++) sT[(m * 125 + p) * SC + sl] = t[m];
    }
  }
  __syncthreads();
  if (!live) return;

#pragma unroll
  for (int ps = 0; ps < NPASS; ps++) {
    const int p = pl + ps * NP;
    if (p < 125) {
      const int k = p / 25, j = (p / 5) % 5, i = p % 5;
      realw f[3];
      tile_divergence<SC>(sT + sl, PL, sG + 25, sG + 50, i, j, k, f);
      realw* a = B.a + (int64_t)gq[ps] * 3 * S + s;
      if (MODE == 2) {
        a[0] = a[0] + (-f[0]);
        a[S] = a[S] + (-f[1]);
        a[2 * S] = a[2 * S] + (-f[2]);
      } else {
        atomicAdd(a, -f[0]);
        atomicAdd(a + S, -f[1]);
        atomicAdd(a + 2 * S, -f[2]);
      }
    }
  }
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_elastic_batch)(hipStream_t st, const BatchArgs& a, int sc, int mode) {
  if (a.nelem <= 0 || a.ngroups <= 0) return;
  const int64_t nb = (int64_t)a.nelem * a.ngroups;
  if (nb > 0x7fffffff) fatal("launch_elastic_batch", "grid too large");
  dim3 grid((unsigned)nb), block(256);
  const GLLk G = g_gll_host;
#define SF_KB(C, M) hipLaunchKernelGGL((k_elastic_batch<C, M>), grid, block, 0, st, a, G)
#define SF_KBM(M)                  \
  if (sc == 2) SF_KB(2, M);        \
  else if (sc == 4) SF_KB(4, M);   \
  else if (sc == 8) SF_KB(8, M);   \
  else fatal("launch_elastic_batch", "shots per workgroup must be 2, 4 or 8");
  if (mode == 2) {
    SF_KBM(2)
  } else {
    SF_KBM(0)
  }
#undef SF_KBM
#undef SF_KB
  SF_LAUNCH_CHECK("k_elastic_batch");
}

}  // namespace sf
