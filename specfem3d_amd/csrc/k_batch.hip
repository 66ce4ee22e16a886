// k_batch.hip -- element stiffness of the batched many-shot engine (elastic, isotropic).
//
// Many independent shots share one mesh. Their wavefields are interleaved shot-fastest: the value
// (global point g, component c, shot s) lives at ((int64_t)g * 3 + c) * S + s, S = padded shot
// count. A workgroup takes one element and a group of SC shots (SC = 4 or 8): the element's ibool,
// moduli and Jacobian terms are loaded once for the SC shots, and workgroups of one element's shot
// groups are adjacent in the grid, so the groups share the L2 lines of each row. The per-point arithmetic is elastic_point.h, the
// same code as the single-shot kernel, so a batched shot reproduces its single-shot run.
//
// Shots are stored in decreasing order of record length; a shot is live at step `it` while
// nstep[s] >= it, and a group whose first shot has finished exits at once (uneven record
// lengths cost nothing after a shot ends). `it` is read from device memory so a captured graph
// of the step can be replayed.
//
// MODE 3 fuses the Newmark update into the coloured stiffness launches. Within a colour no two
// elements share a point, and the colours of a step run in order, so for every global point the
// element of its highest colour (flag BF_LAST) holds the last contribution: it finishes the
// point there -- a *= rmass, v += dt/2 a (corrector), then the next step's predictor
// u += dt v + dt^2/2 a, v += dt/2 a -- because every other element containing the point has
// already gathered u in an earlier colour. The element of the lowest colour (BF_FIRST) writes its
// contribution instead of adding it, so `a` is never cleared except at points that the Stacey and
// source passes add to before the colours (BF_ZERO: cleared by the finisher, never BF_FIRST). A
// point of a single element is gathered, computed and finished in one workgroup with no `a`
// traffic. The sums run in the same order as the separate update path (Stacey, sources, colours
// in order, then the corrector), so the results are the same bits. At a shot's last step the
// finisher applies the corrector only, leaving the state of the last step as the single-shot run
// does.
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

// corrector of the finished step, then (cont) the predictor of the next one; never contracted, so
// the fast build rounds like the separate update kernel (k_update.hip)
__device__ __forceinline__ void newmark_finish(realw& u, realw& v, realw& a, realw M, realw dt, realw dt2, realw dtsq2,
                                               bool cont) {
#pragma clang fp contract(off)
  a = a * M;
  v = v + dt2 * a;
  if (cont) {
    u = u + dt * v + dtsq2 * a;
    v = v + dt2 * a;
  }
}

__device__ __forceinline__ realw& q4(float4& x, int r) { return r == 0 ? x.x : r == 1 ? x.y : r == 2 ? x.z : x.w; }
__device__ __forceinline__ realw c4(const float4& x, int r) { return r == 0 ? x.x : r == 1 ? x.y : r == 2 ? x.z : x.w; }

__device__ __forceinline__ int xcd_remap_b(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// OV = 1 overlays the displacement tile on the stress-product planes (9 instead of 12 planes of
// LDS, one more barrier; the lane keeps its own displacement in registers for the fused update),
// which lets more workgroups share a CU.
template <int NQ, int MODE, int OV>
__global__ void __launch_bounds__(128 * NQ) __attribute__((amdgpu_waves_per_eu(OV ? 4 : 3)))
k_elastic_batch(BatchArgs B, GLLk G) {
  constexpr int SC = 4 * NQ;
  // LDS tile rows (component or stress product, shot) of PST floats, point fastest. PST = 132:
  // the rows of a lane's two quads start 16 banks apart (4 * 132 = 16 mod 32), so the 32 lanes of
  // an LDS pass (16 consecutive points, two quads) read 32 different banks.
  constexpr int PST = 132;
  constexpr int CPL = SC * PST;                   // one component / product plane
  __shared__ realw sBuf[(OV ? 9 : 12) * CPL];
  realw* const sT = sBuf;                         // stress products   [m][s][p]
  // displacement tile [c][s][p]; with OV it shares the planes of the products t3 (m = 6..8),
  // which are written last
  realw* const sU = OV ? sBuf + 6 * CPL : sBuf + 9 * CPL;
  __shared__ realw sG[SF_NGLL_TABLE];                        // hprime, hprimewgll, wgllwgll
  stage_gll(sG, G, threadIdx.x);
  const int it = __hip_atomic_load(B.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int b = xcd_remap_b(blockIdx.x, gridDim.x);
  const int ie = b / B.ngroups;
  const int s0 = (b - ie * B.ngroups) * SC;
  if (B.nstep[s0] < it) return;                   // every shot of this group has finished
  const int q = threadIdx.x % NQ, p = threadIdx.x / NQ;
  const int sq = s0 + 4 * q;                      // first shot of this lane's quad
  const bool live0 = B.nstep[sq] >= it;           // shots sorted: later ones finish first
  bool live[4];
#pragma unroll
  for (int r = 0; r < 4; r++) live[r] = B.nstep[sq + r] >= it;
  const bool active = p < 125;
  const int pp = active ? p : 124;
  const int e = B.elems[ie];
  const int irr = B.irregular[e];
  const int64_t S = B.S;
  const int g = B.ibool[(int64_t)e * 125 + pp];
  const int64_t row = (int64_t)g * 3 * S + sq;
  const int ts = 4 * q;                           // tile slot of the quad
  // Load order matters (the vector-memory counter retires in order): point flags, displacement,
  // moduli and Jacobian terms first, then the lane's accumulator and velocity rows, which are
  // only needed after the element computation and stay in flight across the LDS-only barriers
  // (a: unless this lane writes first; v: only where it finishes the update).
  const unsigned fl = MODE == 3 ? (unsigned)B.flags[(int64_t)e * 125 + pp] : 0u;
  float4 u0 = make_float4(0.f, 0.f, 0.f, 0.f), u1 = u0, u2 = u0;  // staged into LDS below
  if (live0) {
    u0 = *reinterpret_cast<const float4*>(B.u + row);
    u1 = *reinterpret_cast<const float4*>(B.u + row + S);
    u2 = *reinterpret_cast<const float4*>(B.u + row + 2 * S);
  }
  const float2 km = B.kappamu[(int64_t)e * 125 + pp];
  JacPt J = {};
  realw jacl = B.jacobian_regular;
  if (irr) {
    J = load_jac(B.jac, B.jac_plane, (int64_t)(irr - 1) * 125 + pp);
    jacl = J.jac;
  }
  __builtin_amdgcn_sched_barrier(0);
  const bool first = MODE == 3 && (fl & BF_FIRST) != 0;
  const bool last = MODE == 3 && (fl & BF_LAST) != 0;
  const bool mine = active && live0;
  float4 A[3], V[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    A[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    V[c] = A[c];
    if (MODE != 0 && mine && (!first || !live[3])) A[c] = *reinterpret_cast<const float4*>(B.a + row + c * S);
    if (last && mine) V[c] = *reinterpret_cast<const float4*>(B.v + row + c * S);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (active) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      sU[(0 * SC + ts + r) * PST + p] = c4(u0, r);
      sU[(1 * SC + ts + r) * PST + p] = c4(u1, r);
      sU[(2 * SC + ts + r) * PST + p] = c4(u2, r);
    }
  }
  lds_barrier();

  const int k = pp / 25, j = (pp / 5) % 5, i = pp % 5;
  realw t3[4][3];                                 // products t3 of each shot (written last with OV)
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const realw* su = sU + (ts + r) * PST;
    const Grad9 tg = tile_gradient<1>(su, su + SC * PST, su + 2 * SC * PST, gll_hp(sG, irr != 0), i, j, k);
    const Du9 du = map_gradient(tg, irr != 0, J, B.xix_regular);
    const Sigma6 sg = iso_stress(du, km.x, km.y);
    realw t[9];
    stress_products(sg, irr != 0, J, jacl, B.xix_regular, t);
    if (active) {
#pragma unroll
      for (int m = 0; m < (OV ? 6 : 9); m++) sT[(m * SC + ts + r) * PST + p] = t[m];
    }
    t3[r][0] = t[6]; t3[r][1] = t[7]; t3[r][2] = t[8];
  }
  if (OV) {
    lds_barrier();                                // every lane is done with the displacement tile
    if (active) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int m = 6; m < 9; m++) sT[(m * SC + ts + r) * PST + p] = t3[r][m - 6];
    }
  }
  lds_barrier();
  if (!mine) return;

  // one force component at a time for the lane's four shots (keeps few values live)
  realw* a = B.a + row;
  const realw* hpw = gll_hpw(sG, irr != 0);
  const realw* wxy = sG + 50;
  bool cont[4], ends = false;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    cont[r] = live[r] && B.nstep[sq + r] > it;
    ends = ends || (live[r] && !cont[r]);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    realw f[4];
#pragma unroll
    for (int r = 0; r < 4; r++) f[r] = tile_divergence_c<1>(sT + (ts + r) * PST, CPL, hpw, wxy, i, j, k, c);
    if (MODE == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (live[r]) atomicAdd(a + c * S + r, -f[r]);
      continue;
    }
    float4 Ac = A[c];
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (live[r]) q4(Ac, r) = (first ? 0.f : q4(Ac, r)) + (-f[r]);
    if (MODE == 2 || !last) {
      *reinterpret_cast<float4*>(a + c * S) = Ac;
      continue;
    }
    // fused mode, last colour of this point: finish the Newmark update of component c
    const realw M = B.rmass[(int64_t)g * 3 + c];
    float4 Vc = V[c];
    // this step's displacement: the LDS tile without overlay, else the row again (still in cache)
    float4 U;
    if (OV) {
      U = *reinterpret_cast<const float4*>(B.u + row + c * S);
    } else {
      const realw* su = sU + (c * SC + ts) * PST + p;
      U = make_float4(su[0], su[PST], su[2 * PST], su[3 * PST]);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (!live[r]) continue;
      newmark_finish(q4(U, r), q4(Vc, r), q4(Ac, r), M, B.dt, B.dt2, B.dtsq2, cont[r]);
      if (cont[r]) q4(Ac, r) = 0.f;
    }
    if (cont[0]) *reinterpret_cast<float4*>(B.u + row + c * S) = U;
    *reinterpret_cast<float4*>(B.v + row + c * S) = Vc;
    if ((fl & BF_ZERO) || ends) *reinterpret_cast<float4*>(a + c * S) = Ac;
  }
}

// Viscoelastic (3-SLS) variant: each shot also carries its memory variables and stored strain
// (one block per element and shot: plane, point). The tile holds u and dt*v (6 rows per shot; they share the
// LDS of the stress products t2, t3, written after a barrier). Per shot, the arithmetic and its
// order are the single-shot point kernel's (k_elastic_pt): strain at u + dt v, isotropic stress
// minus the memory-variable sums, memory update, stress products. Assembly as k_elastic_batch
// (MODE 2 coloured read-modify-write, 3 coloured with the Newmark update fused).
// Per-shot state is one block per (element, shot), point-fastest within each plane: the lanes of a
// wave that read one shot read consecutive words (a shot-fastest row would spread a wave's reads
// of one shot over S times as many cache lines), and the planes of a block sit at fixed offsets
// from one base address.
__device__ __forceinline__ int64_t rb_idx(int64_t e, int plane, int p, int64_t S, int s) {
  return ((e * S + s) * 18 + plane) * 125 + p;
}
__device__ __forceinline__ int64_t eb_idx(int64_t e, int c, int p, int64_t S, int s) {
  return ((e * S + s) * 6 + c) * 125 + p;
}

template <int NQ, int MODE>
__global__ void __launch_bounds__(128 * NQ) k_visco_batch(BatchArgs B, GLLk G) {
  constexpr int SC = 4 * NQ;
  constexpr int PST = 132;                        // see k_elastic_batch
  constexpr int CPL = SC * PST;
  __shared__ realw sBuf[9 * CPL];
  realw* const sT = sBuf;                         // stress products [m][s][p]
  realw* const sU = sBuf + 3 * CPL;               // u (rows 0..2) and dt*v (rows 3..5), [c][s][p]
  __shared__ realw sG[SF_NGLL_TABLE];
  stage_gll(sG, G, threadIdx.x);
  const int it = __hip_atomic_load(B.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int b = xcd_remap_b(blockIdx.x, gridDim.x);
  const int ie = b / B.ngroups;
  const int s0 = (b - ie * B.ngroups) * SC;
  if (B.nstep[s0] < it) return;
  const int q = threadIdx.x % NQ, p = threadIdx.x / NQ;
  const int sq = s0 + 4 * q;
  const bool live0 = B.nstep[sq] >= it;
  bool live[4];
#pragma unroll
  for (int r = 0; r < 4; r++) live[r] = B.nstep[sq + r] >= it;
  const bool active = p < 125;
  const int pp = active ? p : 124;
  const int e = B.elems[ie];
  const int irr = B.irregular[e];
  const int64_t S = B.S;
  const int64_t ep = (int64_t)e * 125 + pp;
  const int g = B.ibool[ep];
  const int64_t row = (int64_t)g * 3 * S + sq;
  const int ts = 4 * q;
  float4 u[3], V[3];                              // staged into the tile below
#pragma unroll
  for (int c = 0; c < 3; c++) {
    u[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    V[c] = u[c];
    if (live0) {
      u[c] = *reinterpret_cast<const float4*>(B.u + row + c * S);
      V[c] = *reinterpret_cast<const float4*>(B.v + row + c * S);
    }
  }
  const unsigned fl = MODE == 3 ? (unsigned)B.flags[ep] : 0u;
  const bool first = MODE == 3 && (fl & BF_FIRST) != 0;
  const bool last = MODE == 3 && (fl & BF_LAST) != 0;
  const bool mine = active && live0;
  JacPt J = {};
  realw jacl = B.jacobian_regular;
  if (irr) {
    J = load_jac(B.jac, B.jac_plane, (int64_t)(irr - 1) * 125 + pp);
    jacl = J.jac;
  }
  const float2 km = B.kappamu[ep];
  realw fcv[3], fckv[3];
  if (B.fc_index) {
    const int rrow = B.fc_index[ep];
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = B.fc_table[s * B.fc_rows + rrow];
      fckv[s] = B.fc_table[(3 + s) * B.fc_rows + rrow];
    }
  } else {
#pragma unroll
    for (int s = 0; s < 3; s++) {
      fcv[s] = B.fc[s * B.nloc + ep];
      fckv[s] = B.fck[s * B.nloc + ep];
    }
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sU[(c * SC + ts + r) * PST + p] = c4(u[c], r);
        sU[((3 + c) * SC + ts + r) * PST + p] = B.dt * c4(V[c], r);
      }
  }
  // memory variables and stored strain of the first shot, in flight with the tile gather; each
  // later shot's are loaded while the previous shot is computed (two register sets)
  realw Rc[18], Ec[6];
#pragma unroll
  for (int pl = 0; pl < 18; pl++) Rc[pl] = B.R[rb_idx(e, pl, pp, S, sq)];
#pragma unroll
  for (int cc = 0; cc < 6; cc++) Ec[cc] = B.eps[eb_idx(e, cc, pp, S, sq)];
  lds_barrier();

  const int k = pp / 25, j = (pp / 5) % 5, i = pp % 5;
  const realw ONE_THIRD = (realw)1 / (realw)3;
  const realw kappal = km.x, mul = km.y;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int s = sq + r;
    realw Rn[18], En[6];
    if (r < 3) {
#pragma unroll
      for (int pl = 0; pl < 18; pl++) Rn[pl] = B.R[rb_idx(e, pl, pp, S, s + 1)];
#pragma unroll
      for (int cc = 0; cc < 6; cc++) En[cc] = B.eps[eb_idx(e, cc, pp, S, s + 1)];
    }
    __builtin_amdgcn_sched_barrier(0);            // the next shot's loads are issued first
    const realw* su = sU + (ts + r) * PST;
    const Grad9 tg = tile_gradient<1>(su, su + CPL, su + 2 * CPL, gll_hp(sG, irr != 0), i, j, k);
    const Du9 du = map_gradient(tg, irr != 0, J, B.xix_regular);
    const Du9 st = map_gradient(tile_gradient_continue<1>(tg, su + 3 * CPL, su + 4 * CPL, su + 5 * CPL, gll_hp(sG, irr != 0), i, j, k),
                                irr != 0, J, B.xix_regular);
    const realw e_tr = (st.xx + st.yy + st.zz);
    const realw templ = ONE_THIRD * (st.xx + st.yy + st.zz);
    const realw Snp1[6] = {st.xx - templ, st.yy - templ, (realw)0.5 * (st.xy + st.yx), (realw)0.5 * (st.zx + st.xz),
                           (realw)0.5 * (st.zy + st.yz), e_tr};
    Sigma6 sg = iso_stress(du, kappal, mul);
    realw Rv[6][3], Sn[6];
#pragma unroll
    for (int cc = 0; cc < 6; cc++) {
#pragma unroll
      for (int sl = 0; sl < 3; sl++) Rv[cc][sl] = Rc[cc * 3 + sl];
      Sn[cc] = Ec[cc];
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
    if (mine && live[r]) {
#pragma unroll
      for (int sl = 0; sl < 3; sl++) {
        const realw flk = kappal * fckv[sl];
        B.R[rb_idx(e, 5 * 3 + sl, pp, S, s)] =
            B.alpha[sl] * Rv[5][sl] + flk * (B.beta[sl] * Sn[5] + B.gamma[sl] * Snp1[5]);
        const realw flm = mul * fcv[sl];
#pragma unroll
        for (int cc = 0; cc < 5; cc++)
          B.R[rb_idx(e, cc * 3 + sl, pp, S, s)] =
              B.alpha[sl] * Rv[cc][sl] + flm * (B.beta[sl] * Sn[cc] + B.gamma[sl] * Snp1[cc]);
      }
#pragma unroll
      for (int cc = 0; cc < 6; cc++) B.eps[eb_idx(e, cc, pp, S, s)] = Snp1[cc];
    }
    realw t[9];
    stress_products(sg, irr != 0, J, jacl, B.xix_regular, t);
    if (active) {
#pragma unroll
      for (int m = 0; m < 3; m++) sT[(m * SC + ts + r) * PST + p] = t[m];
    }
    // products t2, t3 overlay this shot's u / dt*v rows: written once every lane is past them
    lds_barrier();
    if (active) {
#pragma unroll
      for (int m = 3; m < 9; m++) sT[(m * SC + ts + r) * PST + p] = t[m];
    }
    if (r < 3) {
#pragma unroll
      for (int pl = 0; pl < 18; pl++) Rc[pl] = Rn[pl];
#pragma unroll
      for (int cc = 0; cc < 6; cc++) Ec[cc] = En[cc];
    }
    __builtin_amdgcn_sched_barrier(0);            // at most two shots' state in registers
  }
  lds_barrier();
  if (!mine) return;

  realw* a = B.a + row;
  const realw* hpw = gll_hpw(sG, irr != 0);
  const realw* wxy = sG + 50;
  bool cont[4], ends = false;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    cont[r] = live[r] && B.nstep[sq + r] > it;
    ends = ends || (live[r] && !cont[r]);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    realw f[4];
#pragma unroll
    for (int r = 0; r < 4; r++) f[r] = tile_divergence_c<1>(sT + (ts + r) * PST, CPL, hpw, wxy, i, j, k, c);
    float4 Ac = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!first || !live[3]) Ac = *reinterpret_cast<const float4*>(a + c * S);
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (live[r]) q4(Ac, r) = (first ? 0.f : q4(Ac, r)) + (-f[r]);
    if (MODE == 2 || !last) {
      *reinterpret_cast<float4*>(a + c * S) = Ac;
      continue;
    }
    const realw M = B.rmass[(int64_t)g * 3 + c];
    // this step's u and v rows again (read at the start for the tile; kept out of registers)
    float4 Vc = *reinterpret_cast<const float4*>(B.v + row + c * S);
    float4 U = *reinterpret_cast<const float4*>(B.u + row + c * S);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (!live[r]) continue;
      newmark_finish(q4(U, r), q4(Vc, r), q4(Ac, r), M, B.dt, B.dt2, B.dtsq2, cont[r]);
      if (cont[r]) q4(Ac, r) = 0.f;
    }
    if (cont[0]) *reinterpret_cast<float4*>(B.u + row + c * S) = U;
    *reinterpret_cast<float4*>(B.v + row + c * S) = Vc;
    if ((fl & BF_ZERO) || ends) *reinterpret_cast<float4*>(a + c * S) = Ac;
  }
}

// Batched acoustic stiffness (compute_forces_acoustic.F90:35-726; the arithmetic of k_acoustic per
// point and shot): chi'' -= div((1/rho) grad chi) for SC = 4 NQ shots of one element per
// workgroup. Lane = (GLL point p, quad of 4 consecutive shots) as in k_elastic_batch; the
// potential tile and the three products (rho^-1 jacobian grad chi . grad xi, eta, gamma) go through
// LDS, one row of PST floats per (plane, shot). Every sum runs in k_acoustic's order (the gamma
// contraction reads the same values the column kernel holds in registers), so a batched shot
// reproduces the single-shot run with the same assembly. MODE 0 atomic scatter, 2 plain
// read-modify-write (the launch holds one colour).
template <int NQ, int MODE>
__global__ void __launch_bounds__(128 * NQ) k_acoustic_batch(BatchAcArgs B, GLLk G) {
  constexpr int SC = 4 * NQ;
  constexpr int PST = 132;
  constexpr int CPL = SC * PST;
  __shared__ realw sBuf[4 * CPL];                 // [0] potential tile, [1..3] products t1, t2, t3
  __shared__ realw sG[SF_NGLL_TABLE];
  stage_gll(sG, G, threadIdx.x);
  const int it = __hip_atomic_load(B.it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int b = xcd_remap_b(blockIdx.x, gridDim.x);
  const int ie = b / B.ngroups;
  const int s0 = (b - ie * B.ngroups) * SC;
  if (B.nstep[s0] < it) return;
  const int q = threadIdx.x % NQ, p = threadIdx.x / NQ;
  const int sq = s0 + 4 * q;
  bool live[4];
#pragma unroll
  for (int r = 0; r < 4; r++) live[r] = B.nstep[sq + r] >= it;
  const bool active = p < 125;
  const int pp = active ? p : 124;
  const int e = B.elems[ie];
  const int irr = B.irregular[e];
  const int64_t S = B.S;
  const int g = B.ibool[(int64_t)e * 125 + pp];
  const int64_t row = ((int64_t)g - B.g0) * S + sq;
  const int ts = 4 * q;
  float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (live[0]) c0 = *reinterpret_cast<const float4*>(B.pot + row);
  const realw rho_invl = (realw)1 / B.rho[(int64_t)e * 125 + pp];
  JacPt J = {};
  if (irr) J = load_jac(B.jac, B.jac_plane, (int64_t)(irr - 1) * 125 + pp);
  if (active) {
#pragma unroll
    for (int r = 0; r < 4; r++) sBuf[(ts + r) * PST + p] = c4(c0, r);
  }
  lds_barrier();
  const int k = pp / 25, j = (pp / 5) % 5, i = pp % 5;
  const int b1 = 5 * j + 25 * k, b2 = i + 25 * k, b3 = i + 5 * j;
  const realw* hp = sG;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const realw* s = sBuf + (ts + r) * PST;
    realw a1 = hp[i * 5] * s[b1];
    realw a2 = hp[j * 5] * s[b2];
    realw a3 = hp[k * 5] * s[b3];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      a1 = a1 + hp[i * 5 + l] * s[b1 + l];
      a2 = a2 + hp[j * 5 + l] * s[b2 + 5 * l];
      a3 = a3 + hp[k * 5 + l] * s[b3 + 25 * l];
    }
    realw t1, t2, t3;
    if (irr) {
      const realw dpdx = J.xix * a1 + J.etax * a2 + J.gax * a3;
      const realw dpdy = J.xiy * a1 + J.etay * a2 + J.gay * a3;
      const realw dpdz = J.xiz * a1 + J.etaz * a2 + J.gaz * a3;
      t1 = rho_invl * J.jac * (J.xix * dpdx + J.xiy * dpdy + J.xiz * dpdz);
      t2 = rho_invl * J.jac * (J.etax * dpdx + J.etay * dpdy + J.etaz * dpdz);
      t3 = rho_invl * J.jac * (J.gax * dpdx + J.gay * dpdy + J.gaz * dpdz);
    } else {
      const realw jl = B.jacobian_regular, xr = B.xix_regular;
      t1 = rho_invl * jl * xr * xr * a1;
      t2 = rho_invl * jl * xr * xr * a2;
      t3 = rho_invl * jl * xr * xr * a3;
    }
    if (active) {
      sBuf[(1 * SC + ts + r) * PST + p] = t1;
      sBuf[(2 * SC + ts + r) * PST + p] = t2;
      sBuf[(3 * SC + ts + r) * PST + p] = t3;
    }
  }
  lds_barrier();
  if (!active || !live[0]) return;
  const realw* hpw = sG + 25;
  const realw* wxy = sG + 50;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (MODE == 2) acc = *reinterpret_cast<const float4*>(B.potdd + row);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const realw* s1 = sBuf + (1 * SC + ts + r) * PST;
    const realw* s2 = sBuf + (2 * SC + ts + r) * PST;
    const realw* s3 = sBuf + (3 * SC + ts + r) * PST;
    realw n1 = hpw[i] * s1[b1];
    realw n2 = hpw[j] * s2[b2];
    realw n3 = hpw[k] * s3[b3];
#pragma unroll
    for (int l = 1; l < 5; l++) {
      n1 = n1 + hpw[l * 5 + i] * s1[b1 + l];
      n2 = n2 + hpw[l * 5 + j] * s2[b2 + 5 * l];
      n3 = n3 + hpw[l * 5 + k] * s3[b3 + 25 * l];
    }
    const realw f = wxy[j * 5 + k] * n1 + wxy[i * 5 + k] * n2 + wxy[i * 5 + j] * n3;
    if (!live[r]) continue;
    if (MODE == 0) atomicAdd(B.potdd + row + r, -f);
    else q4(acc, r) = q4(acc, r) + (-f);
  }
  if (MODE == 2) *reinterpret_cast<float4*>(B.potdd + row) = acc;
}

}  // namespace SF_VNS
using namespace SF_VNS;

void SFX(launch_elastic_batch)(hipStream_t st, const BatchArgs& a, int sc, int mode, int overlay) {
  if (a.nelem <= 0 || a.ngroups <= 0) return;
  const int64_t nb = (int64_t)a.nelem * a.ngroups;
  if (nb > 0x7fffffff) fatal("launch_elastic_batch", "grid too large");
  GLLk G = g_gll_host;
  if (SF_FOLD) fold_gll(G, a.xix_regular, a.jacobian_regular);
  dim3 grid((unsigned)nb);
#define SF_KB(NQ, M)                                                                           \
  if (overlay) hipLaunchKernelGGL((k_elastic_batch<NQ, M, 1>), grid, dim3(128 * NQ), 0, st, a, G); \
  else hipLaunchKernelGGL((k_elastic_batch<NQ, M, 0>), grid, dim3(128 * NQ), 0, st, a, G)
#define SF_KBM(M)                  \
  if (sc == 4) SF_KB(1, M);        \
  else if (sc == 8) SF_KB(2, M);   \
  else fatal("launch_elastic_batch", "shots per workgroup must be 4 or 8");
  if (mode == 3) {
    if (!a.v || !a.rmass || !a.flags) fatal("launch_elastic_batch", "fused update needs v, rmass and point flags");
    SF_KBM(3)
  } else if (mode == 2) {
    SF_KBM(2)
  } else {
    SF_KBM(0)
  }
#undef SF_KBM
#undef SF_KB
  SF_LAUNCH_CHECK("k_elastic_batch");
}

void SFX(launch_visco_batch)(hipStream_t st, const BatchArgs& a, int sc, int mode) {
  if (a.nelem <= 0 || a.ngroups <= 0) return;
  const int64_t nb = (int64_t)a.nelem * a.ngroups;
  if (nb > 0x7fffffff) fatal("launch_visco_batch", "grid too large");
  if (!a.R || !a.eps || !a.v) fatal("launch_visco_batch", "memory variables, strain and velocity are needed");
  if (mode == 3 && (!a.rmass || !a.flags)) fatal("launch_visco_batch", "fused update needs rmass and point flags");
  dim3 grid((unsigned)nb);
  GLLk G = g_gll_host;
  if (SF_FOLD) fold_gll(G, a.xix_regular, a.jacobian_regular);
#define SF_KV(NQ, M) hipLaunchKernelGGL((k_visco_batch<NQ, M>), grid, dim3(128 * NQ), 0, st, a, G)
  if (mode == 3) {
    if (sc == 4) SF_KV(1, 3);
    else if (sc == 8) SF_KV(2, 3);
    else fatal("launch_visco_batch", "shots per workgroup must be 4 or 8");
  } else if (mode == 2) {
    if (sc == 4) SF_KV(1, 2);
    else if (sc == 8) SF_KV(2, 2);
    else fatal("launch_visco_batch", "shots per workgroup must be 4 or 8");
  } else {
    fatal("launch_visco_batch", "viscoelastic batches use coloured assembly (2 or 3)");
  }
#undef SF_KV
  SF_LAUNCH_CHECK("k_visco_batch");
}

void SFX(launch_acoustic_batch)(hipStream_t st, const BatchAcArgs& a, int sc, int mode) {
  if (a.nelem <= 0 || a.ngroups <= 0) return;
  const int64_t nb = (int64_t)a.nelem * a.ngroups;
  if (nb > 0x7fffffff) fatal("launch_acoustic_batch", "grid too large");
  dim3 grid((unsigned)nb);
  const GLLk G = g_gll_host;
#define SF_KA(NQ, M) hipLaunchKernelGGL((k_acoustic_batch<NQ, M>), grid, dim3(128 * NQ), 0, st, a, G)
  if (sc != 4 && sc != 8) fatal("launch_acoustic_batch", "shots per workgroup must be 4 or 8");
  if (mode == 0) {
    if (sc == 4) SF_KA(1, 0);
    else SF_KA(2, 0);
  } else {
    if (sc == 4) SF_KA(1, 2);
    else SF_KA(2, 2);
  }
#undef SF_KA
  SF_LAUNCH_CHECK("k_acoustic_batch");
}

}  // namespace sf
