// kernels.h -- launchers of the device kernels (all asynchronous on the given stream).
#pragma once

#include "mesh.h"

namespace sf {

// GLL constants passed to the element kernels by value (kernel-argument segment, read through
// the scalar cache): hp[i*5+l] = l'_l(x_i), hpw[i*5+l] = hp[i*5+l] w_i, wxy[a*5+b] = w_a w_b.
// Attenuation state of the elastic elements, in per-element blocks: the 18 memory-variable planes
// (component-major, SLS second) and the 6 strain planes of one element are contiguous, 125 points
// per plane. A stiffness launch in element order then reads and writes the state as a few
// sequential streams, and each plane of an element is one coalesced 500-byte run.
__host__ __device__ __forceinline__ int64_t r_idx(int64_t e, int plane, int p) { return (e * 18 + plane) * 125 + p; }
__host__ __device__ __forceinline__ int64_t eps_idx(int64_t e, int c, int p) { return (e * 6 + c) * 125 + p; }

struct GLLk {
  realw hp[25];
  realw hpw[25];
  realw wxy[25];
  // FMA builds, regular elements: hp * xix_regular and hpw * (jacobian_regular * xix_regular), which
  // fold the regular mapping and the jacobian weight into the contractions (fold_gll)
  realw hpr[25];
  realw hpwr[25];
};
inline void fold_gll(GLLk& G, realw xix_regular, realw jacobian_regular) {
  const realw jx = jacobian_regular * xix_regular;
  for (int q = 0; q < 25; q++) {
    G.hpr[q] = G.hp[q] * xix_regular;
    G.hpwr[q] = G.hpw[q] * jx;
  }
}
extern GLLk g_gll_host;
void set_gll_constants(const realw* hprime, const realw* hprimewgll, const realw* wgllwgll);

// Newmark pieces (k_update.hip)
void launch_predictor(hipStream_t s, int64_t n, realw* u, realw* v, realw* a, realw dt, realw dt2, realw dtsq2);
void launch_corrector(hipStream_t s, int64_t n, realw* v, realw* a, const realw* rmass, realw dt2);
// with `old`, also PML_displ_old = u + t1 v + t2 a of the corrected field (before the predictor)
void launch_corrector_predictor(hipStream_t s, int64_t n, realw* u, realw* v, realw* a, const realw* rmass,
                                realw dt2_prev, realw dt, realw dt2, realw dtsq2, realw* old = nullptr,
                                realw t1 = 0.f, realw t2 = 0.f);

// element stiffness (k_elastic.hip / k_acoustic.hip)
struct ElasticArgs {
  int nelem;
  const int* elems;
  const int* ibool;
  const int* irregular;
  realw xix_regular, jacobian_regular;
  const realw* jac;        // 10 planes
  int64_t jac_plane;       // nspec_irr*125
  const float2* kappamu;
  const realw* displ;
  const realw* veloc;
  realw* accel;
  realw deltat;
  realw* R;                // 18 planes of nloc
  realw* eps;              // 6 planes of nloc
  const realw* fc;         // 3 planes
  const realw* fck;        // 3 planes
  const realw* fc_table;   // or: 6 planes of fc_rows distinct factor rows (fc 0..2, fck 3..5) ...
  const uint16_t* fc_index;  // ... and a row index per local point (null: use the planes)
  int fc_rows;
  realw* eps_tr3;          // nloc or null
  int64_t nloc;            // nspec*125
  realw alpha[3], beta[3], gamma[3];
  realw* local_out;        // ordered assembly: 3 planes of nloc
};
// mode: 0 atomic scatter, 1 store local contributions (ordered gather), 2 plain read-modify-write
// (the element list is one colour: its elements share no global point)
// att_variant: scheduling of the attenuation state in the point kernel (k_elastic_pt's AV)
void launch_elastic(hipStream_t s, const ElasticArgs& a, bool strain, bool att, bool exact, int mode,
                    int att_variant);
// deviatoric strain planes 0..4 and trace/3 of A.displ; with_trace also writes plane 5 = 3 * trace/3
void launch_element_strain(hipStream_t s, const ElasticArgs& a, realw* eps, realw* eps_tr3, bool with_trace,
                           bool exact);

// C-PML elements (k_pml.hip): stiffness with the convolution memory variables of
// compute_forces_viscoelastic_PML; per C-PML element blocks of 125-point planes
constexpr int SF_PML_NSTATE = 48;
constexpr int SF_PML_NCOEF = 42;
struct PmlArgs {
  int nelem;
  const int* elems;        // element ids ...
  const int* cpml;         // ... and their 0-based C-PML index
  const int* ibool;
  const int* irregular;
  realw xix_regular, jacobian_regular;
  const realw* jac;        // 10 planes
  int64_t jac_plane;
  const float2* kappamu;
  const realw* displ;
  const realw* veloc;
  realw* accel;
  const realw* pml_old;    // 3*nglob
  // per C-PML element: [2 ic] bit m set if coefficient plane m is one value over the element,
  // [2 ic + 1] bit m set if state plane m is live (its multiplier plane is not all zero; dead planes
  // are not stored and enter as zero)
  const uint64_t* masks;
  const int64_t* off;      // [2 ic]: offset of the element's packed coefficient planes, [2 ic + 1]: of its state
  realw* state;            // live state planes, 125 values each, packed per element in plane order
  const realw* coef;       // coefficient planes packed per element in plane order: 1 value if uniform, else 125
  realw t1, t2;            // deltatover2 (1 - 2 theta), deltatsqover2 (1 - theta)
  realw* local_out;        // ordered assembly: 3 planes of nloc
  int64_t nloc;
};
// mode: 0 atomic scatter, 1 store local contributions (ordered gather)
void launch_elastic_pml(hipStream_t s, const PmlArgs& a, bool exact, int mode);
// PML_displ_old at the listed points: u + t1 v + t2 a, with the pending corrector (a *= rmass,
// v += dt2p a) applied first when dt2p_pending != 0
void launch_pml_store_old(hipStream_t s, int n, const int* pts, const realw* u, const realw* v, const realw* a,
                          const realw* rmass, int pending, realw dt2p, realw t1, realw t2, realw* old);
// zero u, v, a at the listed points (Dirichlet condition of the C-PML outer faces)
void launch_zero_points(hipStream_t s, int n, const int* pts, realw* u, realw* v, realw* a);

// batched many-shot stiffness (k_batch.hip): fields interleaved shot-fastest,
// value (g, c, s) at ((int64_t)g * 3 + c) * S + s
struct BatchArgs {
  int nelem;
  const int* elems;
  int ngroups;             // S / shots-per-workgroup
  int S;                   // padded shot count (row stride)
  const int* ibool;
  const int* irregular;
  realw xix_regular, jacobian_regular;
  const realw* jac;
  int64_t jac_plane;
  const float2* kappamu;
  realw* u;
  realw* a;
  const int* nstep;        // per shot slot, decreasing (padding slots 0)
  const int* it;           // device step counter, 1-based current step
  // fused Newmark update (mode 3 only)
  realw* v;
  const realw* rmass;      // inverse mass, 3 per global point
  const unsigned char* flags;  // per local point e*125+p: BF_FIRST | BF_LAST | BF_ZERO
  realw dt, dt2, dtsq2;
  // 3-SLS attenuation (k_visco_batch): per-shot memory variables and strain, one block per
  // (element, shot): value (e, shot, plane, p) at ((e * S + shot) * 18 + plane) * 125 + p
  // (strain: 6 planes)
  realw* R;
  realw* eps;
  const uint16_t* fc_index;    // SLS factor row per local point (or null: fc / fck planes)
  const realw* fc_table;       // 6 planes of fc_rows
  int fc_rows;
  const realw* fc;
  const realw* fck;
  int64_t nloc;
  realw alpha[3], beta[3], gamma[3];
};
enum : unsigned char { BF_FIRST = 1, BF_LAST = 2, BF_ZERO = 4 };
// sc = shots per workgroup (4, 8); mode 0 atomic scatter, 2 plain read-modify-write (one colour),
// 3 one colour with the Newmark update fused (see k_batch.hip); overlay 1 shares the LDS of the
// displacement tile and the stress products
void launch_elastic_batch(hipStream_t s, const BatchArgs& a, int sc, int mode, int overlay, bool exact);
// viscoelastic (3 SLS) batched stiffness; mode 2 or 3
void launch_visco_batch(hipStream_t s, const BatchArgs& a, int sc, int mode, bool exact);

// batched acoustic stiffness (k_batch.hip): potentials interleaved shot-fastest over the acoustic
// range of global points, value (g, s) at (g - g0) * S + s
struct BatchAcArgs {
  int nelem;
  const int* elems;
  int ngroups;             // S / shots-per-workgroup
  int S;                   // padded shot count (row stride)
  int64_t g0;              // first global point of the acoustic range
  const int* ibool;
  const int* irregular;
  realw xix_regular, jacobian_regular;
  const realw* jac;
  int64_t jac_plane;
  const realw* rho;        // rhostore, nspec*125
  const realw* pot;
  realw* potdd;
  const int* nstep;        // per shot slot, decreasing (padding slots 0)
  const int* it;           // device step counter, 1-based current step
};
// sc = shots per workgroup (4, 8); mode 0 atomic scatter, 2 plain read-modify-write (one colour)
void launch_acoustic_batch(hipStream_t s, const BatchAcArgs& a, int sc, int mode, bool exact);

struct AcousticArgs {
  int nelem;
  const int* elems;
  const int* ibool;
  const int* irregular;
  realw xix_regular, jacobian_regular;
  const realw* jac;
  int64_t jac_plane;
  const realw* rho;
  const realw* pot;
  realw* potdd;
  realw* local_out;        // ordered assembly: 1 plane of nloc
};
void launch_acoustic(hipStream_t s, const AcousticArgs& a, bool exact, int mode);

// ordered assembly: out[g*ncomp + c] -= contributions in element order
void launch_ordered_gather(hipStream_t s, int nglob, int ncomp, const int* ptr, const int* idx,
                           const realw* local, int64_t nloc, realw* out);

// boundary terms (k_boundary.hip)
// atomic: one launch with float atomics instead of one launch per round (used when the element
// stiffness scatters with atomics, so the sums have no fixed order anyway)
void launch_stacey_elastic(hipStream_t s, const FacePoints& f, const realw* veloc, realw* accel, realw* b_absorb,
                           int mode, bool atomic);
void launch_stacey_elastic_replay(hipStream_t s, const FacePoints& f, const realw* rec, realw* b_accel);
void launch_stacey_acoustic(hipStream_t s, const FacePoints& f, const realw* potd, realw* potdd, realw* b_absorb,
                            int mode, bool atomic);
void launch_stacey_acoustic_replay(hipStream_t s, const FacePoints& f, const realw* rec, realw* b_potdd);
void launch_coupling_ac_el(hipStream_t s, const FacePoints& f, const realw* displ, realw* potdd, int flip,
                           bool atomic);
void launch_coupling_el_ac(hipStream_t s, const FacePoints& f, const realw* potdd, realw* accel, int flip,
                           bool atomic);
void launch_free_surface(hipStream_t s, const FacePoints& f, realw* pot, realw* potd, realw* potdd);

// sources / receivers (k_sources.hip); FacePoints.ispec holds the owning source / receiver
void launch_add_sources_el(hipStream_t s, const FacePoints& f, const realw* arr, const realw* stf, realw* accel);
void launch_add_sources_ac(hipStream_t s, const FacePoints& f, const realw* arr, const realw* kappa,
                           const realw* stf, realw* potdd);
void launch_adjoint_sources_el(hipStream_t s, const FacePoints& f, const realw* h, const realw* adj, realw* accel);
void launch_adjoint_sources_ac(hipStream_t s, const FacePoints& f, const realw* h, const realw* adj, realw* potdd);
// receivers: `list` holds the local receiver indices to record (n of them; null: receivers 0..n-1);
// out[3 r + c] is written for each listed receiver r
void launch_seismo_el(hipStream_t s, int n, const int* list, const int* iglob, const double* h, const double* nu,
                      const realw* field, realw* out);
// receivers in acoustic elements: pressure = 1 records -sum h field into out[3 r] (out[3 r + 1..2] = 0),
// pressure = 0 the vector (1/rho) grad(field) interpolated and rotated by nu
struct AcReceiverArgs {
  int n;
  const int* list;
  const int* rec_ispec;
  const int* iglob;        // nrec*125 global points of the receiver elements
  const double* h;         // nrec*125 interpolation weights
  const double* nu;        // nrec*9
  const int* irregular;
  realw xix_regular;
  const realw* jac;
  int64_t jac_plane;
  const realw* rho;        // rhostore, nloc
};
void launch_seismo_ac(hipStream_t s, const AcReceiverArgs& a, const realw* field, int pressure, realw* out);

// kernels / norms (k_kernels.hip)
void launch_compute_kernels_el(hipStream_t s, int nelem, const int* elems, const int* ibool, const realw* accel,
                               const realw* b_displ, const realw* eps, const realw* eps_tr3, const realw* b_eps,
                               const realw* b_eps_tr3, int64_t nloc, realw deltat, realw* rho_kl, realw* mu_kl,
                               realw* kappa_kl);
void launch_compute_kernels_ac(hipStream_t s, int nelem, const int* elems, const int* ibool, const int* irregular,
                               realw xix_regular, const realw* jac, int64_t jac_plane, const realw* rhostore,
                               const realw* pot, const realw* b_pot, const realw* b_potdd, realw deltat,
                               realw* rho_ac_kl, realw* kappa_ac_kl);
void launch_max_norm(hipStream_t s, int64_t npts, int ncomp, const realw* f, realw* out);

}  // namespace sf
