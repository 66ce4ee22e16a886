// abi_step.hip -- per-time-step entry points of the Fortran-callable C-ABI, plus transfers.
//
// Call order within a step is fixed by the host (iterate_time.F90, the *_calling routines);
// every entry point enqueues on the mesh's stream and returns (only device->host transfers and
// norms synchronise). FORWARD_OR_ADJOINT: 0 = both fields, 1 = forward (or adjoint) field,
// 3 = backward/reconstructed field.
//
// kernel_3 (mass division + corrector) is deferred: the next update_displacement on the same
// field runs it fused with the predictor (one pass over u, v, a instead of two). Any entry
// point that reads v or a (or writes the field) runs a pending corrector first, so results
// are identical to the eager order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "kernels.h"

namespace sf {
realw* stage_to_device(Mesh* m, const realw* vals, int n, realw* dst);
realw* stf_device_buffer(Mesh* m, int n);
const std::vector<int>& host_ibool(Mesh* m);

static inline int field_of(int foa) { return foa == 3 ? 1 : 0; }

// face-point terms in one atomic launch when the element stiffness scatters with atomics (its
// sums have no fixed order then), else one launch per round in the reference's face order. The
// elastic Stacey term never uses atomics: one launch walks each point's occurrences in round order
// (k_stacey_el_chain, the rounds' sums bit for bit). On the 96x96x48 workload its 0.63 M points take
// 21.7 us so, against 37 us in three rounds and 65 us in one atomic launch; the fluid Stacey term
// and both coupling directions were faster in one atomic launch (4.7 / 15.1 / 5.5 us against
// 18.9 / 19.2 / 18.7 us in rounds).
static inline bool face_atomics(const Mesh* m) { return m->assembly == 0; }

static void flush_el(Mesh* m, int f) {
  if (m->pending_k3_el[f]) {
    const int64_t o = 3 * m->el_g0;  // only the elastic range changes (see Mesh::el_g0)
    launch_corrector(m->stream, 3 * m->el_n, m->el[f].veloc + o, m->el[f].accel + o, m->rmass + o,
                     m->pending_dt2_el[f]);
    m->pending_k3_el[f] = 0;
  }
}
static void flush_ac(Mesh* m, int f) {
  if (m->pending_k3_ac[f]) {
    const int64_t g0 = m->ac_g0;
    launch_corrector(m->stream, m->ac_n, m->ac[f].potd + g0, m->ac[f].potdd + g0, m->rmass_ac + g0,
                     m->pending_dt2_ac[f]);
    m->pending_k3_ac[f] = 0;
  }
}
void flush_all(Mesh* m) {
  for (int f = 0; f < 2; f++) {
    flush_el(m, f);
    flush_ac(m, f);
  }
}

static ElasticArgs elastic_args(Mesh* m, int f, realw deltat) {
  ElasticArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nelem = m->nspec_elastic;
  // atomic launches may walk the elements in the host's locality order; the ordered assembly keeps
  // the processing list (its gather sums in that order)
  a.elems = (m->elastic_launch && m->assembly != 1) ? m->elastic_launch : m->elastic_elems;
  a.ibool = m->ibool;
  a.irregular = m->irregular;
  a.xix_regular = m->xix_regular;
  a.jacobian_regular = m->jacobian_regular;
  a.jac = m->jac;
  a.jac_plane = (int64_t)m->nspec_irregular * 125;
  a.kappamu = m->kappamu;
  a.displ = m->el[f].displ;
  a.veloc = m->el[f].veloc;
  a.accel = m->el[f].accel;
  a.deltat = deltat;
  a.R = m->att[f].R;
  a.eps = m->att[f].eps;
  a.fc = m->factor_common;
  a.fck = m->factor_common_kappa;
  a.fc_table = m->fc_table;
  a.fc_index = m->fc_index;
  a.fc_rows = m->fc_rows;
  a.eps_tr3 = m->simulation_type == 3 ? m->eps_trace_over_3[f] : nullptr;
  a.nloc = (int64_t)m->nspec * 125;
  for (int s = 0; s < 3; s++) {
    a.alpha[s] = f ? m->b_alphaval[s] : m->alphaval[s];
    a.beta[s] = f ? m->b_betaval[s] : m->betaval[s];
    a.gamma[s] = f ? m->b_gammaval[s] : m->gammaval[s];
  }
  a.local_out = m->asm_buf;
  return a;
}

void elastic_forces(Mesh* m, int f, realw deltat, bool strain, bool att) {
  ElasticArgs a = elastic_args(m, f, deltat);
  const bool ordered = m->assembly == 1 && m->asm_ready;
  const bool colored = m->assembly == 2 && m->el_color_elems;
  const bool pml_call = f == 0 && m->pml_n > 0;
  const bool whole = m->timing_scope == 0;
  const bool timed = m->timing_count < m->timing_max && (whole || pml_call);
  if (timed && whole) SF_CHECK(hipEventRecord(m->timing_ev[2 * m->timing_count], m->stream));
  if (colored) {
    for (int c = 0; c < m->el_ncolor; c++) {
      a.elems = m->el_color_elems + m->el_color_off[c];
      a.nelem = m->el_color_off[c + 1] - m->el_color_off[c];
      launch_elastic(m->stream, a, strain, att, m->exact != 0, 2, m->att_variant);
    }
  } else {
    launch_elastic(m->stream, a, strain, att, m->exact != 0, ordered ? 1 : 0, m->att_variant);
  }
  if (pml_call) {
    // C-PML elements (pml_kernel_2_impl of compute_forces_viscoelastic_cuda), after the others: their
    // contributions are summed after those of the stiffness elements
    PmlArgs P;
    std::memset(&P, 0, sizeof(P));
    P.nelem = m->pml_n;
    P.elems = m->pml_elems;
    P.cpml = m->pml_cpml;
    P.ibool = m->ibool;
    P.irregular = m->irregular;
    P.xix_regular = m->xix_regular;
    P.jacobian_regular = m->jacobian_regular;
    P.jac = m->jac;
    P.jac_plane = (int64_t)m->nspec_irregular * 125;
    P.kappamu = m->kappamu;
    P.displ = m->el[0].displ;
    P.veloc = m->el[0].veloc;
    P.accel = m->el[0].accel;
    P.pml_old = m->pml_old;
    P.state = m->pml_state;
    P.coef = m->pml_coef;
    P.masks = m->pml_masks;
    P.off = m->pml_off;
    P.t1 = m->pml_t1;
    P.t2 = m->pml_t2;
    P.local_out = m->asm_buf;
    P.nloc = a.nloc;
    // one launch, float atomics also under coloured assembly: split into its colours (about 1100
    // elements each on the reference's C-PML example) the launches were too small to fill the chip
    // (8 x 35 us against 240 us for one atomic launch before the dead-plane masks)
    if (timed && !whole) SF_CHECK(hipEventRecord(m->timing_ev[2 * m->timing_count], m->stream));
    launch_elastic_pml(m->stream, P, m->exact != 0, ordered ? 1 : 0);
  }
  if (timed) {
    SF_CHECK(hipEventRecord(m->timing_ev[2 * m->timing_count + 1], m->stream));
    m->timing_count++;
  }
  if (ordered)
    launch_ordered_gather(m->stream, m->nglob, 3, m->asm_ptr, m->asm_idx, m->asm_buf, a.nloc, m->el[f].accel);
}

}  // namespace sf

using namespace sf;

static void check_field(Mesh* m, int f, const char* where) {
  if (!m->el[f].displ) fatal(where, f ? "backward elastic field not prepared" : "elastic field not prepared");
}

// ---------------------------------------------------------------- Newmark (update_displacement_cuda.cu)
extern "C" void FC_FUNC_(update_displacement_cuda, UPDATE_DISPLACMENT_CUDA)(long* Mesh_pointer, realw* deltat_F,
                                                                            realw* deltatover2_F,
                                                                            realw* deltatsqover2_F,
                                                                            int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  const int f = field_of(*FORWARD_OR_ADJOINT);
  check_field(m, f, "update_displacement_cuda");
  const int64_t n = 3 * m->el_n, o = 3 * m->el_g0;  // only the elastic range changes
  Wavefield& w = m->el[f];
  const bool pml = f == 0 && m->pml_n > 0;
  if (pml) {
    // update_displacement_cuda.cu:92-124: PML_displ_old from the field before the predictor;
    // PML_displ_new is formed by the C-PML kernel. With a pending corrector the fused
    // corrector-predictor stores it (for every elastic point), else the C-PML points are gathered.
    m->pml_t1 = *deltatover2_F * (1.0f - 2.0f * 0.125f);
    m->pml_t2 = *deltatsqover2_F * (1.0f - 0.125f);
    if (!m->pending_k3_el[f])
      launch_pml_store_old(m->stream, m->pml_npts, m->pml_pts, w.displ, w.veloc, w.accel, m->rmass, 0, 0.f,
                           m->pml_t1, m->pml_t2, m->pml_old);
  }
  if (m->pending_k3_el[f]) {
    launch_corrector_predictor(m->stream, n, w.displ + o, w.veloc + o, w.accel + o, m->rmass + o,
                               m->pending_dt2_el[f], *deltat_F, *deltatover2_F, *deltatsqover2_F,
                               pml ? m->pml_old + o : nullptr, m->pml_t1, m->pml_t2);
    m->pending_k3_el[f] = 0;
  } else {
    launch_predictor(m->stream, n, w.displ + o, w.veloc + o, w.accel + o, *deltat_F, *deltatover2_F,
                     *deltatsqover2_F);
  }
}

extern "C" void FC_FUNC_(update_displacement_ac_cuda, UPDATE_DISPLACEMENT_AC_CUDA)(long* Mesh_pointer,
                                                                                  realw* deltat_F,
                                                                                  realw* deltatover2_F,
                                                                                  realw* deltatsqover2_F,
                                                                                  int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  const int f = field_of(*FORWARD_OR_ADJOINT);
  AcField& a = m->ac[f];
  if (!a.pot) fatal("update_displacement_ac_cuda", "acoustic field not prepared");
  const int64_t g0 = m->ac_g0;  // only the acoustic range changes (see Mesh::ac_g0)
  if (m->pending_k3_ac[f]) {
    launch_corrector_predictor(m->stream, m->ac_n, a.pot + g0, a.potd + g0, a.potdd + g0, m->rmass_ac + g0,
                               m->pending_dt2_ac[f], *deltat_F, *deltatover2_F, *deltatsqover2_F);
    m->pending_k3_ac[f] = 0;
  } else {
    launch_predictor(m->stream, m->ac_n, a.pot + g0, a.potd + g0, a.potdd + g0, *deltat_F, *deltatover2_F,
                     *deltatsqover2_F);
  }
}

extern "C" void FC_FUNC_(kernel_3_a_cuda, KERNEL_3_A_CUDA)(long* Mesh_pointer, realw* deltatover2_F,
                                                           realw* b_deltatover2_F, int* APPROXIMATE_OCEAN_LOAD,
                                                           int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*APPROXIMATE_OCEAN_LOAD) fatal("kernel_3_a_cuda", "ocean load is not supported");
  const int foa = *FORWARD_OR_ADJOINT;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->el[f].displ) continue;
    flush_el(m, f);
    m->pending_k3_el[f] = 1;
    m->pending_dt2_el[f] = f ? *b_deltatover2_F : *deltatover2_F;
  }
}

extern "C" void FC_FUNC_(kernel_3_b_cuda, KERNEL_3_B_CUDA)(long* Mesh_pointer, realw* deltatover2_F,
                                                           realw* b_deltatover2_F, int* FORWARD_OR_ADJOINT) {
  (void)Mesh_pointer; (void)deltatover2_F; (void)b_deltatover2_F; (void)FORWARD_OR_ADJOINT;
  fatal("kernel_3_b_cuda", "ocean load is not supported");
}

extern "C" void FC_FUNC_(kernel_3_acoustic_cuda, KERNEL_3_ACOUSTIC_CUDA)(long* Mesh_pointer, realw* deltatover2_F,
                                                                         realw* b_deltatover2_F,
                                                                         int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  const int foa = *FORWARD_OR_ADJOINT_f;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot) continue;
    flush_ac(m, f);
    m->pending_k3_ac[f] = 1;
    m->pending_dt2_ac[f] = f ? *b_deltatover2_F : *deltatover2_F;
  }
}

// ---------------------------------------------------------------- forces
// compute_forces_viscoelastic_cuda.cu:1115
extern "C" void FC_FUNC_(compute_forces_viscoelastic_cuda, COMPUTE_FORCES_VISCOELASTIC_CUDA)(
    long* Mesh_pointer, int* iphase, realw* deltat, int* nspec_outer_elastic, int* nspec_inner_elastic,
    int* COMPUTE_AND_STORE_STRAIN, int* ATTENUATION, int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  // NPROC=1: every element is "inner"; all elements are computed in phase 2
  (void)nspec_outer_elastic; (void)nspec_inner_elastic;
  if (*iphase != 2) return;
  const bool strain = *COMPUTE_AND_STORE_STRAIN != 0;
  const bool att = *ATTENUATION != 0;
  if (att && !m->att[0].R) fatal("compute_forces_viscoelastic_cuda", "attenuation arrays not prepared");
  const int foa = *FORWARD_OR_ADJOINT_f;
  // FORWARD_OR_ADJOINT = 0 computes the adjoint field, then the backward field
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (foa == 0 && !m->el[1].displ && f == 1) continue;
    check_field(m, f, "compute_forces_viscoelastic_cuda");
    flush_el(m, f);
    if (strain && !m->att[f].eps) fatal("compute_forces_viscoelastic_cuda", "strain arrays not prepared");
    elastic_forces(m, f, *deltat, strain, att && m->att[f].R != nullptr);
  }
}

// compute_forces_acoustic_cuda.cu:238
extern "C" void FC_FUNC_(compute_forces_acoustic_cuda, COMPUTE_FORCES_ACOUSTIC_CUDA)(long* Mesh_pointer, int* iphase,
                                                                                   int* nspec_outer_acoustic,
                                                                                   int* nspec_inner_acoustic,
                                                                                   int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)nspec_outer_acoustic; (void)nspec_inner_acoustic;
  if (*iphase != 2) return;
  const int foa = *FORWARD_OR_ADJOINT_f;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot) continue;
    flush_ac(m, f);
    AcousticArgs a;
    a.nelem = m->nspec_acoustic;
    a.elems = m->acoustic_elems;
    a.ibool = m->ibool;
    a.irregular = m->irregular;
    a.xix_regular = m->xix_regular;
    a.jacobian_regular = m->jacobian_regular;
    a.jac = m->jac;
    a.jac_plane = (int64_t)m->nspec_irregular * 125;
    a.rho = m->rho;
    a.pot = m->ac[f].pot;
    a.potdd = m->ac[f].potdd;
    a.local_out = m->asm_ac_buf;
    const bool ordered = m->assembly == 1 && m->asm_ac_buf;
    if (m->assembly == 2 && m->ac_color_elems) {
      for (int c = 0; c < m->ac_ncolor; c++) {
        a.elems = m->ac_color_elems + m->ac_color_off[c];
        a.nelem = m->ac_color_off[c + 1] - m->ac_color_off[c];
        launch_acoustic(m->stream, a, m->exact != 0, 2);
      }
    } else {
      launch_acoustic(m->stream, a, m->exact != 0, ordered ? 1 : 0);
    }
    if (ordered)
      launch_ordered_gather(m->stream, m->nglob, 1, m->asm_ac_ptr, m->asm_ac_idx, m->asm_ac_buf,
                            (int64_t)m->nspec * 125, m->ac[f].potdd);
  }
}

// compute_forces_acoustic_cuda.cu:334 -- zero chi, chi', chi'' on the fluid free surface. It
// commutes with a deferred acoustic corrector (zeroed points stay zero through it).
extern "C" void FC_FUNC_(acoustic_enforce_free_surf_cuda, ACOUSTIC_ENFORCE_FREE_SURF_CUDA)(
    long* Mesh_pointer, int* ABSORB_INSTEAD_OF_FREE_SURFACE, int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*ABSORB_INSTEAD_OF_FREE_SURFACE) return;
  const int foa = *FORWARD_OR_ADJOINT;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot) continue;
    launch_free_surface(m->stream, m->free_ac, m->ac[f].pot, m->ac[f].potd, m->ac[f].potdd);
  }
}

// ---------------------------------------------------------------- C-PML
// pml_compute_accel_cuda.cu:245 -- Dirichlet condition on the outer faces of the C-PML layers: u, v, a
// set to zero at their points (the reference also zeroes PML_displ_old/new there, which the next
// predictor overwrites before any use)
extern "C" void FC_FUNC_(pml_impose_boundary_condition_elastic_cuda,
                         PML_IMPOSE_BOUNDARY_CONDITION_ELASTIC_CUDA)(long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->pml_conditions || m->pml_nbc == 0) return;
  flush_el(m, 0);
  launch_zero_points(m->stream, m->pml_nbc, m->pml_bc_pts, m->el[0].displ, m->el[0].veloc, m->el[0].accel);
}

// transfer_fields_cuda.cu -- PML_displ_old/new(NDIM, NGLL3, NSPEC_CPML). The library keeps
// PML_displ_old per global point and forms PML_displ_new from u and v inside the C-PML kernel, so
// "from" returns the stored old values and new = u + deltatover2 (1 - 2 theta) v + deltatsqover2
// (1 - theta) a of the current field (the reference's array between the predictor and the
// stiffness); "to" sets the old values (equal at shared points) and ignores new, which the next
// predictor recomputes in the reference as well.
static void pml_displ_check(Mesh* m, int* size, const char* where) {
  if (!m->pml_state) fatal(where, "C-PML arrays not prepared");
  if (*size != 3 * 125 * m->nspec_cpml) fatal(where, "size must be NDIM*NGLL3*NSPEC_CPML");
}

extern "C" void FC_FUNC_(transfer_pml_displ_from_device, TRANSFER_PML_DISPL_FROM_DEVICE)(int* size,
                                                                                       realw* PML_displ_old,
                                                                                       realw* PML_displ_new,
                                                                                       long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  pml_displ_check(m, size, "transfer_pml_displ_from_device");
  flush_el(m, 0);
  const size_t n3 = 3 * (size_t)m->nglob;
  std::vector<realw> u(n3), v(n3), a(n3), od(n3);
  SF_CHECK(hipMemcpyAsync(u.data(), m->el[0].displ, sizeof(realw) * n3, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipMemcpyAsync(v.data(), m->el[0].veloc, sizeof(realw) * n3, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipMemcpyAsync(a.data(), m->el[0].accel, sizeof(realw) * n3, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipMemcpyAsync(od.data(), m->pml_old, sizeof(realw) * n3, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
  const std::vector<int>& ib = host_ibool(m);
  std::vector<int> el(m->pml_n), cp(m->pml_n);
  SF_CHECK(hipMemcpy(el.data(), m->pml_elems, sizeof(int) * el.size(), hipMemcpyDeviceToHost));
  SF_CHECK(hipMemcpy(cp.data(), m->pml_cpml, sizeof(int) * cp.size(), hipMemcpyDeviceToHost));
  for (int q = 0; q < m->pml_n; q++)
    for (int p = 0; p < 125; p++) {
      const int64_t g = ib[(int64_t)el[q] * 125 + p];
      for (int c = 0; c < 3; c++) {
        const int64_t d = ((int64_t)cp[q] * 125 + p) * 3 + c;
        PML_displ_old[d] = od[3 * g + c];
        PML_displ_new[d] = u[3 * g + c] + m->pml_t1 * v[3 * g + c] + m->pml_t2 * a[3 * g + c];
      }
    }
}

extern "C" void FC_FUNC_(transfer_pml_displ_to_device, TRANSFER_PML_DISPL_TO_DEVICE)(int* size, realw* PML_displ_old,
                                                                                   realw* PML_displ_new,
                                                                                   long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)PML_displ_new;
  pml_displ_check(m, size, "transfer_pml_displ_to_device");
  const size_t n3 = 3 * (size_t)m->nglob;
  std::vector<realw> od(n3);
  SF_CHECK(hipMemcpyAsync(od.data(), m->pml_old, sizeof(realw) * n3, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
  const std::vector<int>& ib = host_ibool(m);
  std::vector<int> el(m->pml_n), cp(m->pml_n);
  SF_CHECK(hipMemcpy(el.data(), m->pml_elems, sizeof(int) * el.size(), hipMemcpyDeviceToHost));
  SF_CHECK(hipMemcpy(cp.data(), m->pml_cpml, sizeof(int) * cp.size(), hipMemcpyDeviceToHost));
  for (int q = 0; q < m->pml_n; q++)
    for (int p = 0; p < 125; p++) {
      const int64_t g = ib[(int64_t)el[q] * 125 + p];
      for (int c = 0; c < 3; c++) od[3 * g + c] = PML_displ_old[((int64_t)cp[q] * 125 + p) * 3 + c];
    }
  SF_CHECK(hipMemcpyAsync(m->pml_old, od.data(), sizeof(realw) * n3, hipMemcpyHostToDevice, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
}

// ---------------------------------------------------------------- boundaries
// compute_stacey_viscoelastic_cuda.cu:39. Forward field: apply (and, when saving the forward
// run, fill b_absorb_field(3,25,nfaces) for the host to write). Backward field of a kernel
// run: subtract the record the host read into b_absorb_field.
extern "C" void FC_FUNC_(compute_stacey_viscoelastic_cuda, COMPUTE_STACEY_VISCOELASTIC_CUDA)(
    long* Mesh_pointer, int* iphasef, realw* b_absorb_field, int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || !m->stacey || m->abs_el.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT_f;
  const int64_t rec_len = 3 * 25 * (int64_t)m->num_abs_boundary_faces;
  if (foa == 0 || foa == 1) {
    flush_el(m, 0);
    const bool store = m->simulation_type == 1 && m->save_forward && b_absorb_field;
    launch_stacey_elastic(m->stream, m->abs_el, m->el[0].veloc, m->el[0].accel, m->b_absorb_field, store ? 1 : 0,
                          false);
    if (store) {
      SF_CHECK(hipMemcpyAsync(b_absorb_field, m->b_absorb_field, sizeof(realw) * rec_len, hipMemcpyDeviceToHost,
                              m->stream));
      SF_CHECK(hipStreamSynchronize(m->stream));
    }
  }
  if ((foa == 0 || foa == 3) && m->simulation_type == 3 && m->el[1].displ) {
    flush_el(m, 1);
    SF_CHECK(hipMemcpyAsync(m->b_absorb_field, b_absorb_field, sizeof(realw) * rec_len, hipMemcpyHostToDevice,
                            m->stream));
    launch_stacey_elastic_replay(m->stream, m->abs_el, m->b_absorb_field, m->el[1].accel);
    SF_CHECK(hipStreamSynchronize(m->stream));  // the host buffer is reused for the next record
  }
}

// compute_stacey_viscoelastic_cuda.cu:224 -- undo-attenuation / checkpoint runs step the
// reconstructed field forward in time, so Stacey is applied to it directly.
extern "C" void FC_FUNC_(compute_stacey_viscoelastic_undoatt_cuda, COMPUTE_STACEY_VISCOELASTIC_UNDOATT_CUDA)(
    long* Mesh_pointer, int* iphasef, int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || !m->stacey || m->abs_el.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT_f;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->el[f].displ) continue;
    flush_el(m, f);
    launch_stacey_elastic(m->stream, m->abs_el, m->el[f].veloc, m->el[f].accel, nullptr, 0, false);
  }
}

extern "C" void FC_FUNC_(compute_stacey_acoustic_cuda, COMPUTE_STACEY_ACOUSTIC_CUDA)(long* Mesh_pointer,
                                                                                   int* iphasef,
                                                                                   realw* h_b_absorb_potential,
                                                                                   int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || !m->stacey || m->abs_ac.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT_f;
  const int64_t rec_len = 25 * (int64_t)m->num_abs_boundary_faces;
  if (foa == 0 || foa == 1) {
    flush_ac(m, 0);
    const bool store = m->simulation_type == 1 && m->save_forward && h_b_absorb_potential;
    launch_stacey_acoustic(m->stream, m->abs_ac, m->ac[0].potd, m->ac[0].potdd, m->b_absorb_potential, store ? 1 : 0,
                           face_atomics(m));
    if (store) {
      SF_CHECK(hipMemcpyAsync(h_b_absorb_potential, m->b_absorb_potential, sizeof(realw) * rec_len,
                              hipMemcpyDeviceToHost, m->stream));
      SF_CHECK(hipStreamSynchronize(m->stream));
    }
  }
  if ((foa == 0 || foa == 3) && m->simulation_type == 3 && m->ac[1].pot) {
    flush_ac(m, 1);
    SF_CHECK(hipMemcpyAsync(m->b_absorb_potential, h_b_absorb_potential, sizeof(realw) * rec_len,
                            hipMemcpyHostToDevice, m->stream));
    launch_stacey_acoustic_replay(m->stream, m->abs_ac, m->b_absorb_potential, m->ac[1].potdd);
    SF_CHECK(hipStreamSynchronize(m->stream));
  }
}

extern "C" void FC_FUNC_(compute_stacey_acoustic_undoatt_cuda, COMPUTE_STACEY_ACOUSTIC_UNDOATT_CUDA)(
    long* Mesh_pointer, int* iphasef, int* FORWARD_OR_ADJOINT_f) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || !m->stacey || m->abs_ac.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT_f;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot) continue;
    flush_ac(m, f);
    launch_stacey_acoustic(m->stream, m->abs_ac, m->ac[f].potd, m->ac[f].potdd, nullptr, 0, face_atomics(m));
  }
}

// compute_coupling_cuda.cu:39 -- fluid side: chi'' += w u.n (adjoint field: u -> -u)
extern "C" void FC_FUNC_(compute_coupling_ac_el_cuda, COMPUTE_COUPLING_AC_EL_CUDA)(long* Mesh_pointer, int* iphasef,
                                                                                 int* num_coupling_ac_el_facesf,
                                                                                 int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || *num_coupling_ac_el_facesf == 0 || m->coupling.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot || !m->el[f].displ) continue;
    flush_ac(m, f);
    const int flip = (m->simulation_type != 1 && f == 0) ? 1 : 0;
    launch_coupling_ac_el(m->stream, m->coupling, m->el[f].displ, m->ac[f].potdd, flip, face_atomics(m));
  }
}

// compute_coupling_cuda.cu:135 -- solid side: a += w n (-chi'') (adjoint field: sign flip)
extern "C" void FC_FUNC_(compute_coupling_el_ac_cuda, COMPUTE_COUPLING_EL_AC_CUDA)(long* Mesh_pointer, int* iphasef,
                                                                                 int* num_coupling_ac_el_facesf,
                                                                                 int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*iphasef != 1 || *num_coupling_ac_el_facesf == 0 || m->coupling.n == 0) return;
  const int foa = *FORWARD_OR_ADJOINT;
  for (int f = 0; f < 2; f++) {
    if (foa == 1 && f == 1) continue;
    if (foa == 3 && f == 0) continue;
    if (!m->ac[f].pot || !m->el[f].displ) continue;
    flush_ac(m, f);
    flush_el(m, f);
    const int flip = (m->simulation_type != 1 && f == 0) ? 1 : 0;
    launch_coupling_el_ac(m->stream, m->coupling, m->ac[f].potdd, m->el[f].accel, flip, face_atomics(m));
  }
}

// ---------------------------------------------------------------- sources
// compute_add_sources_viscoelastic_cuda.cu:39; stf_used = real(stf) per source (double on the host)
static void add_sources_el(Mesh* m, int f, const double* h_stf, int nsrc) {
  if (m->src_pts.n == 0 || nsrc <= 0) return;
  std::vector<realw> s(nsrc);
  bool any = false;
  for (int i = 0; i < nsrc; i++) {
    s[i] = (realw)h_stf[i];
    any = any || h_stf[i] != 0.0;
  }
  if (!any) return;  // as the reference's GPU wrapper: nothing to add when all STFs are zero
  flush_el(m, f);
  realw* d = stf_device_buffer(m, nsrc);
  stage_to_device(m, s.data(), nsrc, d);
  launch_add_sources_el(m->stream, m->src_pts, m->src_array, d, m->el[f].accel);
}

extern "C" void FC_FUNC_(compute_add_sources_el_cuda, COMPUTE_ADD_SOURCES_EL_CUDA)(long* Mesh_pointer,
                                                                                 double* h_stf_pre_compute,
                                                                                 int* h_NSOURCES) {
  Mesh* m = from_handle(Mesh_pointer);
  add_sources_el(m, 0, h_stf_pre_compute, *h_NSOURCES);
}

extern "C" void FC_FUNC_(compute_add_sources_el_s3_cuda, COMPUTE_ADD_SOURCES_EL_S3_CUDA)(long* Mesh_pointer,
                                                                                       double* h_stf_pre_compute,
                                                                                       int* h_NSOURCES) {
  Mesh* m = from_handle(Mesh_pointer);
  check_field(m, 1, "compute_add_sources_el_s3_cuda");
  add_sources_el(m, 1, h_stf_pre_compute, *h_NSOURCES);
}

// compute_add_sources_viscoelastic_cuda.cu:224 -- adjoint sources at the receivers for step it:
// h_source_adjoint(3, nadj_rec_local, NTSTEP_BETWEEN_READ_ADJSRC), sample
// NTSTEP_BETWEEN_READ_ADJSRC - mod(it-1, NTSTEP_BETWEEN_READ_ADJSRC)
extern "C" void FC_FUNC_(add_sources_el_sim_type_2_or_3, ADD_SOURCES_EL_SIM_TYPE_2_OR_3)(
    long* Mesh_pointer, realw* h_source_adjoint, int* nrec, int* nadj_rec_local, int* NTSTEP_BETWEEN_READ_ADJSRC,
    int* it) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)nrec;
  const int nadj = *nadj_rec_local;
  if (nadj <= 0 || m->adj_pts.n == 0) return;
  const int nt = *NTSTEP_BETWEEN_READ_ADJSRC;
  const int isample = nt - ((*it - 1) % nt);  // 1-based
  const realw* src = h_source_adjoint + (int64_t)3 * nadj * (isample - 1);
  flush_el(m, 0);
  stage_to_device(m, src, 3 * nadj, m->adj_src);
  launch_adjoint_sources_el(m->stream, m->adj_pts, m->adj_h, m->adj_src, m->el[0].accel);
}

// compute_add_sources_acoustic_cuda.cu: sources in acoustic elements, chi'' -= sourcearray(1) stf / kappa
// (compute_add_sources_acoustic.f90:95-124); stf_used = real(stf) per source
static void add_sources_ac(Mesh* m, int f, const double* h_stf, int nsrc, const char* where) {
  if (m->src_ac_pts.n == 0 || nsrc <= 0) return;
  if (!m->ac[f].pot) fatal(where, f ? "backward acoustic field not prepared" : "acoustic field not prepared");
  if (!m->src_kappa) fatal(where, "prepare_fields_acoustic_device not called");
  if (nsrc < m->nsources) fatal(where, "NSOURCES smaller than the number of prepared sources");
  std::vector<realw> s(m->nsources);
  for (int i = 0; i < m->nsources; i++) s[i] = (realw)h_stf[i];
  flush_ac(m, f);
  realw* d = stf_device_buffer(m, m->nsources);
  stage_to_device(m, s.data(), m->nsources, d);
  launch_add_sources_ac(m->stream, m->src_ac_pts, m->src_array, m->src_kappa, d, m->ac[f].potdd);
}

extern "C" void FC_FUNC_(compute_add_sources_ac_cuda, COMPUTE_ADD_SOURCES_AC_CUDA)(long* Mesh_pointer,
                                                                                 int* NSOURCESf,
                                                                                 double* h_stf_pre_compute,
                                                                                 int* run_number_of_the_source) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)run_number_of_the_source;  // NB_RUNS_ACOUSTIC_GPU = 1 (prepare_constants_device refuses others)
  add_sources_ac(m, 0, h_stf_pre_compute, *NSOURCESf, "compute_add_sources_ac_cuda");
}

extern "C" void FC_FUNC_(compute_add_sources_ac_s3_cuda, COMPUTE_ADD_SOURCES_AC_s3_CUDA)(long* Mesh_pointer,
                                                                                       int* NSOURCESf,
                                                                                       double* h_stf_pre_compute,
                                                                                       int* run_number_of_the_source) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)run_number_of_the_source;
  add_sources_ac(m, 1, h_stf_pre_compute, *NSOURCESf, "compute_add_sources_ac_s3_cuda");
}

// compute_add_sources_acoustic_cuda.cu (add_sources_ac_sim_2_or_3_cuda): adjoint sources at receivers in
// acoustic elements, chi'' += adj(1, irec, isample) hxi heta hgamma with the same sample index as the
// elastic ones
extern "C" void FC_FUNC_(add_sources_ac_sim_2_or_3_cuda, ADD_SOURCES_AC_SIM_2_OR_3_CUDA)(
    long* Mesh_pointer, realw* h_source_adjoint, int* nrec, int* nadj_rec_local, int* NTSTEP_BETWEEN_READ_ADJSRC,
    int* it) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)nrec;
  const int nadj = *nadj_rec_local;
  if (nadj <= 0 || m->adj_ac_pts.n == 0) return;
  if (!m->ac[0].pot) fatal("add_sources_ac_sim_2_or_3_cuda", "acoustic field not prepared");
  const int nt = *NTSTEP_BETWEEN_READ_ADJSRC;
  const int isample = nt - ((*it - 1) % nt);  // 1-based
  const realw* src = h_source_adjoint + (int64_t)3 * nadj * (isample - 1);
  flush_ac(m, 0);
  stage_to_device(m, src, 3 * nadj, m->adj_src);
  launch_adjoint_sources_ac(m->stream, m->adj_ac_pts, m->adj_h, m->adj_src, m->ac[0].potdd);
}

// ---------------------------------------------------------------- seismograms (write_seismograms_cuda.cu:34)
extern "C" void FC_FUNC_(compute_seismograms_cuda, COMPUTE_SEISMOGRAMS_CUDA)(
    long* Mesh_pointer_f, realw* seismograms_d, realw* seismograms_v, realw* seismograms_a, realw* seismograms_p,
    int* seismo_currentf, int* nlength_seismogramf, int* itf, int* it_endf, int* ACOUSTIC_SIMULATION,
    int* ELASTIC_SIMULATION, int* USE_TRICK_FOR_BETTER_PRESSURE) {
  Mesh* m = from_handle(Mesh_pointer_f);
  (void)ACOUSTIC_SIMULATION; (void)ELASTIC_SIMULATION;
  if (m->nrec_local <= 0) return;
  const bool trick = USE_TRICK_FOR_BETTER_PRESSURE && *USE_TRICK_FOR_BETTER_PRESSURE;
  if (m->save_p && m->nrec_el > 0)
    fatal("compute_seismograms_cuda", "pressure seismograms at receivers in elastic elements are not supported");
  const int cur = *seismo_currentf;
  // SIMULATION_TYPE 3 records the reconstructed (backward) field
  const int f = (m->simulation_type == 3 && m->el[1].displ) ? 1 : 0;
  const int fa = (m->simulation_type == 3 && m->ac[1].pot) ? 1 : 0;
  const int64_t off = (int64_t)(cur - 1) * m->nrec_local * 3;
  if (m->nrec_el > 0) {
    if (!m->el[f].displ) fatal("compute_seismograms_cuda", "elastic field not prepared");
    if (m->save_v || m->save_a) flush_el(m, f);
    const Wavefield& w = m->el[f];
    if (m->save_d) launch_seismo_el(m->stream, m->nrec_el, m->rec_el_list, m->rec_iglob, m->rec_h, m->rec_nu, w.displ, m->seismo_d + off);
    if (m->save_v) launch_seismo_el(m->stream, m->nrec_el, m->rec_el_list, m->rec_iglob, m->rec_h, m->rec_nu, w.veloc, m->seismo_v + off);
    if (m->save_a) launch_seismo_el(m->stream, m->nrec_el, m->rec_el_list, m->rec_iglob, m->rec_h, m->rec_nu, w.accel, m->seismo_a + off);
  }
  if (m->nrec_ac > 0) {
    // receivers in fluid elements (write_seismograms_cuda.cu: seismotype 1-3 from the gradient of the
    // potentials, 4 = pressure from the potential with the pressure trick, else its second derivative)
    const AcField& a = m->ac[fa];
    if (!a.pot) fatal("compute_seismograms_cuda", "acoustic field not prepared");
    if (m->save_v || m->save_a || (m->save_p && !trick)) flush_ac(m, fa);
    AcReceiverArgs ra;
    ra.n = m->nrec_ac;
    ra.list = m->rec_ac_list;
    ra.rec_ispec = m->rec_ispec;
    ra.iglob = m->rec_iglob;
    ra.h = m->rec_h;
    ra.nu = m->rec_nu;
    ra.irregular = m->irregular;
    ra.xix_regular = m->xix_regular;
    ra.jac = m->jac;
    ra.jac_plane = (int64_t)m->nspec_irregular * 125;
    ra.rho = m->rho;
    if (m->save_d) launch_seismo_ac(m->stream, ra, a.pot, 0, m->seismo_d + off);
    if (m->save_v) launch_seismo_ac(m->stream, ra, a.potd, 0, m->seismo_v + off);
    if (m->save_a) launch_seismo_ac(m->stream, ra, a.potdd, 0, m->seismo_a + off);
    if (m->save_p) launch_seismo_ac(m->stream, ra, trick ? a.pot : a.potdd, 1, m->seismo_p + off);
  }
  if (cur == *nlength_seismogramf || *itf == *it_endf) {
    // host arrays seismograms_?(NDIM, nrec_local, nlength); pressure in component 1, the others zero
    const size_t bytes = sizeof(realw) * (size_t)cur * m->nrec_local * 3;
    if (m->save_d) SF_CHECK(hipMemcpyAsync(seismograms_d, m->seismo_d, bytes, hipMemcpyDeviceToHost, m->stream));
    if (m->save_v) SF_CHECK(hipMemcpyAsync(seismograms_v, m->seismo_v, bytes, hipMemcpyDeviceToHost, m->stream));
    if (m->save_a) SF_CHECK(hipMemcpyAsync(seismograms_a, m->seismo_a, bytes, hipMemcpyDeviceToHost, m->stream));
    if (m->save_p) SF_CHECK(hipMemcpyAsync(seismograms_p, m->seismo_p, bytes, hipMemcpyDeviceToHost, m->stream));
    SF_CHECK(hipStreamSynchronize(m->stream));
  }
}

// ---------------------------------------------------------------- kernels (compute_kernels_cuda.cu:39)
extern "C" void FC_FUNC_(compute_kernels_elastic_cuda, COMPUTE_KERNELS_ELASTIC_CUDA)(long* Mesh_pointer,
                                                                                   realw* deltat_f) {
  Mesh* m = from_handle(Mesh_pointer);
  check_field(m, 1, "compute_kernels_elastic_cuda");
  if (!m->rho_kl) fatal("compute_kernels_elastic_cuda", "kernel arrays not prepared");
  flush_el(m, 0);
  flush_el(m, 1);
  const int64_t nl = (int64_t)m->nspec * 125;
  if (m->undo_attenuation) {
    // the reconstructed strain is recomputed from b_displ (compute_kernels_cuda.cu:57-100)
    ElasticArgs a = elastic_args(m, 1, 0.f);
    launch_element_strain(m->stream, a, m->att[1].eps, m->eps_trace_over_3[1], m->attenuation != 0, m->exact != 0);
  }
  launch_compute_kernels_el(m->stream, m->nspec_elastic, m->elastic_elems, m->ibool, m->el[0].accel, m->el[1].displ,
                            m->att[0].eps, m->eps_trace_over_3[0], m->att[1].eps, m->eps_trace_over_3[1], nl,
                            *deltat_f, m->rho_kl, m->mu_kl, m->kappa_kl);
}

// compute_strain_cuda.cu: strain of the (backward) displacement for undo-attenuation runs
extern "C" void FC_FUNC_(compute_strain_cuda, COMPUTE_strain_CUDA)(long* Mesh_pointer, int* FORWARD_OR_ADJOINT) {
  Mesh* m = from_handle(Mesh_pointer);
  const int f = field_of(*FORWARD_OR_ADJOINT);
  check_field(m, f, "compute_strain_cuda");
  if (!m->att[f].eps || !m->eps_trace_over_3[f]) fatal("compute_strain_cuda", "strain arrays not prepared");
  flush_el(m, f);
  ElasticArgs a = elastic_args(m, f, 0.f);
  launch_element_strain(m->stream, a, m->att[f].eps, m->eps_trace_over_3[f], m->attenuation != 0, m->exact != 0);
}

extern "C" void FC_FUNC_(compute_kernels_acoustic_cuda, COMPUTE_KERNELS_ACOUSTIC_CUDA)(long* Mesh_pointer,
                                                                                     realw* deltat_f) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->ac[0].pot || !m->ac[1].pot) fatal("compute_kernels_acoustic_cuda", "acoustic fields not prepared");
  if (!m->rho_ac_kl) fatal("compute_kernels_acoustic_cuda", "acoustic kernel arrays not prepared");
  flush_ac(m, 0);
  flush_ac(m, 1);
  launch_compute_kernels_ac(m->stream, m->nspec_acoustic, m->acoustic_elems, m->ibool, m->irregular, m->xix_regular,
                            m->jac, (int64_t)m->nspec_irregular * 125, m->rho,
                            m->ac_kernel_dd ? m->ac[0].potdd : m->ac[0].pot, m->ac[1].pot,
                            m->ac[1].potdd, *deltat_f, m->rho_ac_kl, m->kappa_ac_kl);
}

// ---------------------------------------------------------------- norms (check_fields_cuda.cu:375-470)
extern "C" void FC_FUNC_(get_norm_elastic_from_device, GET_NORM_ELASTIC_FROM_DEVICE)(realw* norm, long* Mesh_pointer,
                                                                                   int* type) {
  Mesh* m = from_handle(Mesh_pointer);
  const int f = field_of(*type);
  check_field(m, f, "get_norm_elastic_from_device");
  launch_max_norm(m->stream, m->nglob, 3, m->el[f].displ, m->d_norm);
  SF_CHECK(hipMemcpyAsync(m->h_norm, m->d_norm, sizeof(realw), hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
  *norm = m->h_norm[0];
}

extern "C" void FC_FUNC_(get_norm_acoustic_from_device, GET_NORM_ACOUSTIC_FROM_DEVICE)(realw* norm,
                                                                                     long* Mesh_pointer,
                                                                                     int* sim_type) {
  Mesh* m = from_handle(Mesh_pointer);
  const int f = (*sim_type == 3) ? 1 : 0;
  if (!m->ac[f].pot) fatal("get_norm_acoustic_from_device", "acoustic field not prepared");
  launch_max_norm(m->stream, m->nglob, 1, m->ac[f].pot, m->d_norm);
  SF_CHECK(hipMemcpyAsync(m->h_norm, m->d_norm, sizeof(realw), hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
  *norm = m->h_norm[0];
}

// ---------------------------------------------------------------- transfers (transfer_fields_cuda.cu)
static void h2d(Mesh* m, realw* d, const realw* h, int64_t n) {
  SF_CHECK(hipMemcpyAsync(d, h, sizeof(realw) * (size_t)n, hipMemcpyHostToDevice, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
}
static void d2h(Mesh* m, realw* h, const realw* d, int64_t n) {
  SF_CHECK(hipMemcpyAsync(h, d, sizeof(realw) * (size_t)n, hipMemcpyDeviceToHost, m->stream));
  SF_CHECK(hipStreamSynchronize(m->stream));
}

#define SF_XFER3(NAME, F, DIR)                                                                                  \
  extern "C" void FC_FUNC_(NAME, X)(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer) { \
    Mesh* m = from_handle(Mesh_pointer);                                                                      \
    check_field(m, F, #NAME);                                                                                 \
    flush_el(m, F);                                                                                           \
    DIR(m, m->el[F].displ, displ, *size);                                                                     \
    DIR(m, m->el[F].veloc, veloc, *size);                                                                     \
    DIR(m, m->el[F].accel, accel, *size);                                                                     \
  }
#define SF_XFER1(NAME, F, MEMBER, DIR)                                                   \
  extern "C" void FC_FUNC_(NAME, X)(int* size, realw* arr, long* Mesh_pointer) {        \
    Mesh* m = from_handle(Mesh_pointer);                                                 \
    check_field(m, F, #NAME);                                                            \
    flush_el(m, F);                                                                      \
    DIR(m, m->el[F].MEMBER, arr, *size);                                                 \
  }
#define H2D(m, d, h, n) h2d(m, d, h, n)
#define D2H(m, d, h, n) d2h(m, h, d, n)

SF_XFER3(transfer_fields_el_to_device, 0, H2D)
SF_XFER3(transfer_fields_el_from_device, 0, D2H)
SF_XFER3(transfer_b_fields_to_device, 1, H2D)
SF_XFER3(transfer_b_fields_from_device, 1, D2H)
SF_XFER1(transfer_displ_to_device, 0, displ, H2D)
SF_XFER1(transfer_displ_from_device, 0, displ, D2H)
SF_XFER1(transfer_veloc_to_device, 0, veloc, H2D)
SF_XFER1(transfer_veloc_from_device, 0, veloc, D2H)
SF_XFER1(transfer_accel_to_device, 0, accel, H2D)
SF_XFER1(transfer_accel_from_device, 0, accel, D2H)
SF_XFER1(transfer_b_displ_to_device, 1, displ, H2D)
SF_XFER1(transfer_b_displ_from_device, 1, displ, D2H)
SF_XFER1(transfer_b_veloc_to_device, 1, veloc, H2D)
SF_XFER1(transfer_b_veloc_from_device, 1, veloc, D2H)
SF_XFER1(transfer_b_accel_to_device, 1, accel, H2D)
SF_XFER1(transfer_b_accel_from_device, 1, accel, D2H)

#define SF_XFERAC3(NAME, F, DIR)                                                                                 \
  extern "C" void FC_FUNC_(NAME, X)(int* size, realw* p, realw* pd, realw* pdd, long* Mesh_pointer) {           \
    Mesh* m = from_handle(Mesh_pointer);                                                                       \
    if (!m->ac[F].pot) fatal(#NAME, "acoustic field not prepared");                                            \
    flush_ac(m, F);                                                                                            \
    DIR(m, m->ac[F].pot, p, *size);                                                                            \
    DIR(m, m->ac[F].potd, pd, *size);                                                                          \
    DIR(m, m->ac[F].potdd, pdd, *size);                                                                        \
  }
#define SF_XFERAC1(NAME, F, MEMBER, DIR)                                          \
  extern "C" void FC_FUNC_(NAME, X)(int* size, realw* arr, long* Mesh_pointer) { \
    Mesh* m = from_handle(Mesh_pointer);                                          \
    if (!m->ac[F].pot) fatal(#NAME, "acoustic field not prepared");               \
    flush_ac(m, F);                                                               \
    DIR(m, m->ac[F].MEMBER, arr, *size);                                          \
  }
SF_XFERAC3(transfer_fields_ac_to_device, 0, H2D)
SF_XFERAC3(transfer_fields_ac_from_device, 0, D2H)
SF_XFERAC3(transfer_b_fields_ac_to_device, 1, H2D)
SF_XFERAC3(transfer_b_fields_ac_from_device, 1, D2H)
SF_XFERAC1(transfer_potential_ac_from_device, 0, pot, D2H)
SF_XFERAC1(transfer_b_potential_ac_from_device, 1, pot, D2H)
SF_XFERAC1(transfer_b_potential_ac_to_device, 1, pot, H2D)
SF_XFERAC1(transfer_b_potential_dot_dot_ac_from_device, 1, potdd, D2H)
SF_XFERAC1(transfer_b_potential_dot_dot_ac_to_device, 1, potdd, H2D)
SF_XFERAC1(transfer_dot_dot_to_device, 0, potdd, H2D)
SF_XFERAC1(transfer_dot_dot_from_device, 0, potdd, D2H)
SF_XFERAC1(transfer_b_dot_dot_from_device, 1, potdd, D2H)

// transfer_fields_cuda.cu: sensitivity kernels (5,5,5,NSPEC_AB) back to the host
extern "C" void FC_FUNC_(transfer_kernels_el_to_host, TRANSFER_KERNELS_EL_TO_HOST)(long* Mesh_pointer,
                                                                                 realw* h_rho_kl, realw* h_mu_kl,
                                                                                 realw* h_kappa_kl,
                                                                                 realw* h_cijkl_kl, int* NSPEC_AB) {
  Mesh* m = from_handle(Mesh_pointer);
  (void)h_cijkl_kl;
  if (!m->rho_kl) fatal("transfer_kernels_el_to_host", "kernel arrays not prepared");
  const int64_t n = (int64_t)*NSPEC_AB * 125;
  d2h(m, h_rho_kl, m->rho_kl, n);
  d2h(m, h_mu_kl, m->mu_kl, n);
  d2h(m, h_kappa_kl, m->kappa_kl, n);
}

extern "C" void FC_FUNC_(transfer_strain_from_device, TRANSFER_STRAIN_FROM_DEVICE)(
    long* Mesh_pointer, realw* epsilondev_xx, realw* epsilondev_yy, realw* epsilondev_xy, realw* epsilondev_xz,
    realw* epsilondev_yz, realw* epsilondev_trace, int* size_epsilondev) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->att[0].eps) fatal("transfer_strain_from_device", "strain arrays not prepared");
  const int64_t nl = (int64_t)m->nspec * 125, n = std::min<int64_t>(*size_epsilondev, nl);
  realw* dst[6] = {epsilondev_xx, epsilondev_yy, epsilondev_xy, epsilondev_xz, epsilondev_yz, epsilondev_trace};
  std::vector<realw> E(6 * (size_t)nl);
  d2h(m, E.data(), m->att[0].eps, 6 * nl);
  for (int c = 0; c < 6; c++)
    if (dst[c])
      for (int64_t q = 0; q < n; q++) dst[c][q] = E[(size_t)eps_idx(q / 125, c, (int)(q % 125))];
}

extern "C" void FC_FUNC_(transfer_b_strain_to_device, TRANSFER_B_strain_TO_DEVICE)(
    long* Mesh_pointer, realw* b_epsilondev_xx, realw* b_epsilondev_yy, realw* b_epsilondev_xy,
    realw* b_epsilondev_xz, realw* b_epsilondev_yz, realw* b_epsilondev_trace, int* size_epsilondev) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->att[1].eps) fatal("transfer_b_strain_to_device", "backward strain arrays not prepared");
  const int64_t nl = (int64_t)m->nspec * 125, n = std::min<int64_t>(*size_epsilondev, nl);
  realw* src[6] = {b_epsilondev_xx, b_epsilondev_yy, b_epsilondev_xy, b_epsilondev_xz, b_epsilondev_yz,
                   b_epsilondev_trace};
  std::vector<realw> E(6 * (size_t)nl);
  d2h(m, E.data(), m->att[1].eps, 6 * nl);   // components not passed keep their device values
  for (int c = 0; c < 6; c++)
    if (src[c])
      for (int64_t q = 0; q < n; q++) E[(size_t)eps_idx(q / 125, c, (int)(q % 125))] = src[c][q];
  h2d(m, m->att[1].eps, E.data(), 6 * nl);
}

static void rmem_copy(Mesh* m, int f, realw* const* arrs, int64_t nR, bool to_device) {
  if (!m->att[f].R) fatal("transfer_rmemory", "attenuation arrays not prepared");
  const int64_t nl = (int64_t)m->nspec * 125;
  if (nR != 3 * nl) fatal("transfer_rmemory", "size_R must be N_SLS*125*NSPEC");
  std::vector<realw> plane(18 * (size_t)nl);
  if (!to_device) d2h(m, plane.data(), m->att[f].R, 18 * nl);
  for (int c = 0; c < 6; c++)
    for (int64_t q = 0; q < nl; q++)
      for (int s = 0; s < 3; s++) {
        realw& pv = plane[(size_t)r_idx(q / 125, c * 3 + s, (int)(q % 125))];
        if (to_device) pv = arrs[c][3 * q + s];
        else arrs[c][3 * q + s] = pv;
      }
  if (to_device) h2d(m, m->att[f].R, plane.data(), 18 * nl);
}

extern "C" void FC_FUNC_(transfer_rmemory_from_device, TRANSFER_RMEMORY_FROM_DEVICE)(long* Mesh_pointer, realw* R_xx,
                                                                                   realw* R_yy, realw* R_xy,
                                                                                   realw* R_xz, realw* R_yz,
                                                                                   realw* R_trace, int* size_R) {
  Mesh* m = from_handle(Mesh_pointer);
  realw* a[6] = {R_xx, R_yy, R_xy, R_xz, R_yz, R_trace};
  rmem_copy(m, 0, a, *size_R, false);
}

extern "C" void FC_FUNC_(transfer_b_rmemory_to_device, TRANSFER_B_RMEMORY_TO_DEVICE)(long* Mesh_pointer,
                                                                                   realw* b_R_xx, realw* b_R_yy,
                                                                                   realw* b_R_xy, realw* b_R_xz,
                                                                                   realw* b_R_yz, realw* b_R_trace,
                                                                                   int* size_R) {
  Mesh* m = from_handle(Mesh_pointer);
  realw* a[6] = {b_R_xx, b_R_yy, b_R_xy, b_R_xz, b_R_yz, b_R_trace};
  rmem_copy(m, 1, a, *size_R, true);
}

extern "C" void FC_FUNC_(transfer_kernels_ac_to_host, TRANSFER_KERNELS_AC_TO_HOST)(long* Mesh_pointer,
                                                                                 realw* h_rho_ac_kl,
                                                                                 realw* h_kappa_ac_kl,
                                                                                 int* NSPEC_AB) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->rho_ac_kl) fatal("transfer_kernels_ac_to_host", "acoustic kernel arrays not prepared");
  const int64_t n = (int64_t)*NSPEC_AB * 125;
  d2h(m, h_rho_ac_kl, m->rho_ac_kl, n);
  d2h(m, h_kappa_ac_kl, m->kappa_ac_kl, n);
}

// ---------------------------------------------------------------- library extensions (no reference counterpart)
// sf_set_options: exact = 1 runs the element kernels without fused multiply-add (the CPU
// reference's rounding); assembly = 1 sums element contributions in element order, 2 colours.
extern "C" void sf_set_options(long* Mesh_pointer, int exact, int assembly);
extern "C" void sf_synchronize(long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  flush_all(m);
  SF_CHECK(hipStreamSynchronize(m->stream));
}

// Time the next `max_launches` element-stiffness launches with HIP events on the mesh stream.
extern "C" void sf_timing_start(long* Mesh_pointer, int max_launches) {
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  if (m->timing_ev) {
    for (int i = 0; i < 2 * m->timing_max; i++) (void)hipEventDestroy(m->timing_ev[i]);
    delete[] m->timing_ev;
  }
  m->timing_max = max_launches > 0 ? max_launches : 0;
  m->timing_count = 0;
  m->timing_ev = m->timing_max ? new hipEvent_t[2 * m->timing_max] : nullptr;
  for (int i = 0; i < 2 * m->timing_max; i++) SF_CHECK(hipEventCreate(&m->timing_ev[i]));
}

// Total milliseconds and number of timed launches since sf_timing_start (synchronises).
extern "C" void sf_timing_read(long* Mesh_pointer, double* total_ms, int* count) {
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  double t = 0.0;
  for (int i = 0; i < m->timing_count; i++) {
    float ms = 0.f;
    SF_CHECK(hipEventElapsedTime(&ms, m->timing_ev[2 * i], m->timing_ev[2 * i + 1]));
    t += ms;
  }
  *total_ms = t;
  *count = m->timing_count;
}
