// mesh.h -- the opaque handle behind the reference's `long* Mesh_pointer`.
//
// Device data layout (chosen for gfx950, not copied from the reference's Mesh struct):
//   * wavefields displ/veloc/accel: float[3*nglob], xyz interleaved (same as Fortran
//     displ(3,NGLOB), so host transfers are plain copies);
//   * ibool: int32[nspec*125] 0-based, i fastest (Fortran ibool(i,j,k,ispec) order);
//   * kappa/mu: one float2 per local point so each point needs one 8-byte load;
//   * irregular Jacobians: 10 SoA planes [t][nspec_irr*125] (xix..gammaz, jacobian) so a
//     wavefront's 64 lanes read 64 consecutive floats of one plane;
//   * attenuation memory variables and stored strain: per-element blocks, the 18 memory-variable
//     planes (comp*3 + sls, comp = xx,yy,xy,xz,yz,trace) of one element contiguous, then the next
//     element: value (e, plane, p) at (e*18 + plane)*125 + p (r_idx in kernels.h); strain likewise
//     with 6 planes xx,yy,xy,xz,yz,trace at (e*6 + c)*125 + p (eps_idx);
//   * inverse mass: float[3*nglob] interleaved (rmassx,rmassy,rmassz) so the corrector is a
//     pure float4 stream over the flat 3*nglob arrays;
//   * boundary faces are flattened to face points with the global index, normal, weight and
//     the per-point impedances gathered once at prepare time.
#pragma once

#include <vector>

#include "common.h"

namespace sf {

struct Wavefield {
  realw* displ = nullptr;
  realw* veloc = nullptr;
  realw* accel = nullptr;
};

struct AcField {
  realw* pot = nullptr;
  realw* potd = nullptr;
  realw* potdd = nullptr;
};

struct AttField {            // memory variables + stored strain of one elastic wavefield
  realw* R = nullptr;        // 18 * nloc: per-element blocks of 18 planes (r_idx in kernels.h)
  realw* eps = nullptr;      // 6 * nloc: per-element blocks of 6 planes xx,yy,xy,xz,yz,trace (eps_idx)
};

struct FacePoints {          // flattened boundary/coupling face points
  int n = 0;                 // number of face points (faces * 25)
  int* iglob = nullptr;
  int* ispec = nullptr;      // element of the face point (0-based)
  int* ploc = nullptr;       // local point index p = i + 5j + 25k
  realw* normal = nullptr;   // 3*n
  realw* jacw = nullptr;     // n
  realw* imp_p = nullptr;    // rho*vp (elastic Stacey) or 1/(rho*c) factor (acoustic)
  realw* imp_s = nullptr;    // rho*vs (elastic Stacey)
  // Points are sorted into "rounds": round r holds the r-th occurrence of each global point in
  // the host's face order, so one launch per round can update without atomics and the sum at
  // shared edge/corner points keeps the reference's face order.
  int nrounds = 0;
  int round_off[17] = {0};
  int* slot = nullptr;       // original host slot (face*25 + igll) of each point, for records
  int* next = nullptr;       // index of the same global point's next occurrence (next round), -1 after the last
};

struct Snapshot {            // one saved state of a wavefield (save_forward_arrays_undoatt)
  realw* el = nullptr;       // displ, veloc, accel (3 x 3*nglob)
  realw* ac = nullptr;       // potential, first and second derivatives (3 x nglob)
  realw* R = nullptr;        // 18 memory-variable planes (attenuation)
};

struct Mesh {
  int nspec = 0, nglob = 0, nspec_irregular = 0;
  int simulation_type = 1;
  int save_forward = 0;
  int undo_attenuation = 0;        // UNDO_ATTENUATION_AND_OR_PML: checkpointed kernel runs
  int myrank = 0;

  // geometry
  int* ibool = nullptr;
  int* irregular = nullptr;
  realw xix_regular = 0.f, jacobian_regular = 0.f;
  realw* jac = nullptr;            // 10 planes of nspec_irr*125
  int jac_from_host = 0;           // 1 if the determinant came from the host (sf_set_jacobian)
  int* h_ibool = nullptr;          // host copy (0-based) for host-side index work

  // domains
  int nspec_elastic = 0, nspec_acoustic = 0;
  int* elastic_elems = nullptr;    // element lists in processing order
  int* elastic_launch = nullptr;   // optional locality order of the elastic elements for the atomic
                                   // and coloured launches (sf_set_element_order); null: elastic_elems
  int* acoustic_elems = nullptr;
  unsigned char* h_is_elastic = nullptr;
  unsigned char* h_is_acoustic = nullptr;

  // elastic material and mass
  float2* kappamu = nullptr;
  realw* rho = nullptr;            // rhostore (acoustic forces, kernels)
  realw* kappa_ac = nullptr;       // kappastore for acoustic elements
  realw* rmass = nullptr;          // 3*nglob inverse mass (elastic)
  realw* rmass_ac = nullptr;       // nglob inverse mass (acoustic)
  int64_t el_g0 = 0, el_n = 0;     // elastic points lie in [el_g0, el_g0 + el_n) (el_g0 a multiple of 4);
                                   // the elastic Newmark streams cover only this range
  int64_t ac_g0 = 0, ac_n = 0;     // acoustic points lie in [ac_g0, ac_g0 + ac_n) (ac_g0 a multiple of 4);
                                   // the acoustic Newmark streams cover only this range

  // wavefields: [0] forward/adjoint, [1] backward/reconstructed
  Wavefield el[2];
  AcField ac[2];
  int compute_and_store_strain = 0;
  int attenuation = 0;
  AttField att[2];
  realw* factor_common = nullptr;  // 3 planes (mu)
  realw* factor_common_kappa = nullptr;  // 3 planes
  // the same factors as a table of distinct (factor_common[3], factor_common_kappa[3]) rows and a
  // 16-bit row index per local point (when at most 65535 distinct rows exist; the factors depend
  // only on the point's Q values, which are few): 2 B per point instead of 24 B in the kernel
  realw* fc_table = nullptr;       // 6 planes of fc_rows
  uint16_t* fc_index = nullptr;    // nloc
  int fc_rows = 0;
  realw alphaval[N_SLS], betaval[N_SLS], gammaval[N_SLS];
  realw b_alphaval[N_SLS], b_betaval[N_SLS], b_gammaval[N_SLS];
  realw* eps_trace_over_3[2] = {nullptr, nullptr};

  // boundaries
  int stacey = 0;
  FacePoints abs_el, abs_ac;       // Stacey face points in elastic / acoustic elements
  realw* b_absorb_field = nullptr; // device buffer for one step of stored tractions (3*abs_el.n)
  realw* b_absorb_potential = nullptr;
  FacePoints free_ac;              // acoustic free-surface points
  FacePoints coupling;             // fluid-solid coupling points (on acoustic elements)
  int num_abs_boundary_faces = 0;
  int* h_abs_face_point_el = nullptr;  // host map from record slot (face*25+igll) to abs_el index

  // C-PML (elastic; prepare_fields_elastic_pml): the stiffness kernel skips the C-PML elements and
  // a second kernel computes them with their convolution memory variables
  int pml_conditions = 0;          // PML_CONDITIONS of prepare_constants_device
  int nspec_cpml = 0;
  int pml_n = 0;                   // listed C-PML elements (elastic ones, processing order) ...
  int* pml_elems = nullptr;        // ... their element ids ...
  int* pml_cpml = nullptr;         // ... and their 0-based C-PML index
  // per C-PML element, packed planes: the live state planes (of 48: 27 three-term strain memory
  // variables, 12 one-term ones, 9 displacement memory variables), 125 values each, and the
  // coefficient planes (42: alpha 9, beta 9, abar 5, strain 18, wgll_cube * rho), one value for a
  // plane uniform over the element, else 125
  realw* pml_state = nullptr;
  realw* pml_coef = nullptr;
  uint64_t* pml_masks = nullptr;   // 2 per C-PML element: uniform coefficient planes, live state planes
  int64_t* pml_off = nullptr;      // 2 per C-PML element: offsets (floats) of its coefficient and state planes
  // PML_displ_old at every global point (3*nglob; it is a function of the point's u, v, a before the
  // predictor); PML_displ_new is recomputed in the kernel from u and v after the predictor (a = 0)
  realw* pml_old = nullptr;
  int pml_npts = 0;                // global points of the C-PML elements (store list)
  int* pml_pts = nullptr;
  int pml_nbc = 0;                 // global points on the outer faces of the layers (Dirichlet list)
  int* pml_bc_pts = nullptr;
  realw pml_t1 = 0.f, pml_t2 = 0.f;  // deltatover2 (1 - 2 theta), deltatsqover2 (1 - theta)

  // sources
  int nsources = 0;
  FacePoints src_pts;              // (source, point) pairs of sources in elastic elements
  FacePoints src_ac_pts;           // (source, point) pairs of sources in acoustic elements
  realw* src_array = nullptr;      // nsrc*125*3
  unsigned char* h_src_elastic = nullptr;
  unsigned char* h_src_acoustic = nullptr;
  realw* src_kappa = nullptr;      // nsrc*125 kappastore at the source element's points (acoustic sources)

  // receivers
  int nrec_local = 0;
  int* rec_iglob = nullptr;        // nrec*125
  double* rec_h = nullptr;         // nrec*125 products hxi*heta*hgamma (double)
  double* rec_nu = nullptr;        // nrec*9
  int nlength_seismogram = 0;
  realw* seismo_d = nullptr;       // device buffers [nlength][nrec][3]
  realw* seismo_v = nullptr;
  realw* seismo_a = nullptr;
  realw* seismo_p = nullptr;
  int save_d = 0, save_v = 0, save_a = 0, save_p = 0;
  unsigned char* h_rec_elastic = nullptr;
  int* h_rec_ispec = nullptr;
  int* rec_ispec = nullptr;        // device: element of each local receiver (0-based)
  int nrec_el = 0, nrec_ac = 0;    // local receivers in elastic / acoustic elements ...
  int* rec_el_list = nullptr;      // ... and their local receiver indices (device)
  int* rec_ac_list = nullptr;

  // adjoint sources (SIMULATION_TYPE 2/3)
  int nadj_rec_local = 0;
  int ntstep_between_read_adjsrc = 0;
  realw* adj_h = nullptr;          // nadj*125 float products (CUSTOM_REAL hlagrange)
  FacePoints adj_pts;              // (adjoint receiver, point) pairs in elastic elements
  FacePoints adj_ac_pts;           // (adjoint receiver, point) pairs in acoustic elements
  realw* adj_src = nullptr;        // one step of adjoint sources (3*nadj)
  unsigned char* h_adj_elastic = nullptr;

  // sensitivity kernels
  realw* rho_kl = nullptr;
  realw* mu_kl = nullptr;
  realw* kappa_kl = nullptr;
  realw* rho_ac_kl = nullptr;
  realw* kappa_ac_kl = nullptr;

  // execution
  hipStream_t stream = nullptr;
  // Defaults for a host that never calls sf_set_options (the Fortran drop-in): the FMA build and,
  // once prepare_fields_elastic_device knows whether attenuation is on, coloured assembly without it
  // (float atomics are the limit of the light elastic kernel) and atomic assembly with it (the
  // colour launches re-read more state than the atomics write back).
  int exact = 0;                   // 1: kernels built without fused multiply-add (oracle order)
  int assembly = -1;               // 0 atomic scatter-add, 1 ordered gather (oracle order), 2 coloured,
                                   // -1 automatic (resolved at prepare_fields_elastic_device)
  int att_variant = 4;             // attenuation-state scheduling of the point kernel (0..4)
  int ac_kernel_dd = 0;            // 1: fluid kernels read the adjoint potential_dot_dot (misfit gradient)
  // deferred corrector (kernel_3) per field, fused into the next predictor when possible
  int pending_k3_el[2] = {0, 0};
  realw pending_dt2_el[2] = {0.f, 0.f};
  int pending_k3_ac[2] = {0, 0};
  realw pending_dt2_ac[2] = {0.f, 0.f};
  // ordered-assembly structures (CSR of contributions per global point, sorted by element)
  int* asm_ptr = nullptr;          // nglob+1
  int* asm_idx = nullptr;          // slots e*125+p
  realw* asm_buf = nullptr;        // 3*nspec*125 per-element local forces
  int asm_ready = 0;
  int* asm_ac_ptr = nullptr;       // same for the acoustic elements (1 component)
  int* asm_ac_idx = nullptr;
  realw* asm_ac_buf = nullptr;
  // coloured assembly (assembly = 2): element lists regrouped by colour; elements of one
  // colour share no global point, colour c is elems[off[c] .. off[c+1])
  int el_ncolor = 0, ac_ncolor = 0;
  int el_color_off[33] = {0}, ac_color_off[33] = {0};
  int color_version = 0;           // bumped whenever a colouring is (re)built
  int* el_color_elems = nullptr;
  int* ac_color_elems = nullptr;

  // device-resident checkpoints of checkpointed adjoint runs (abi_checkpoint.hip)
  std::vector<Snapshot> snapshots;
  int history_frames = 0;
  realw* history_el = nullptr;     // history_frames x 3*nglob displacement frames
  realw* history_ac = nullptr;     // history_frames x 2*nglob (potential, second derivative)

  // scratch for reductions
  realw* d_norm = nullptr;
  realw* h_norm = nullptr;         // pinned

  // optional HIP-event timing of the element-stiffness launches (sf_timing_*)
  int timing_max = 0;
  int timing_count = 0;
  int timing_scope = 0;            // 0: the whole elastic stiffness call, 1: its C-PML launch only
  hipEvent_t* timing_ev = nullptr; // 2 * timing_max events (start, stop)
};

Mesh* from_handle(long* p);

}  // namespace sf
