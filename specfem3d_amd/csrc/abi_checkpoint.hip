// abi_checkpoint.hip -- device-resident snapshots and displacement history for checkpointed
// adjoint (sensitivity-kernel) runs.
//
// The reference's UNDO_ATTENUATION_AND_OR_PML kernel run (iterate_time_undoatt.F90:29-632)
//   * writes the forward state at the start of every time subset to a file
//     (save_forward_arrays_undoatt, save_forward_arrays.f90:102-179: potentials, displ, veloc,
//     accel and, with attenuation, the memory variables) and reads them back in reverse order
//     (read_forward_arrays_undoatt, read_forward_arrays.f90:133-229);
//   * recomputes one subset of the forward field and keeps every step's b_displ in a host
//     buffer (b_displ_elastic_buffer), copying one frame back to the device per adjoint step
//     (iterate_time_undoatt.F90:448-520).
// Here both stores live in HBM: a snapshot is a device-to-device copy of the field state and the
// history is an array of displacement frames, so the adjoint loop never crosses PCIe. The strain
// is not part of a snapshot: as in the reference, the host recomputes it from the loaded
// displacement (compute_strain_cuda) after a load.
#include <cstring>

#include "kernels.h"

namespace sf {
void flush_all(Mesh* m);

static inline int field_of_foa(int foa) { return foa == 3 ? 1 : 0; }

static realw* ck_alloc(int64_t n, const char* where) {
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(realw) * (size_t)(n > 0 ? n : 1)) != hipSuccess)
    fatal(where, "out of device memory for checkpoint storage (use a larger NT_DUMP_ATTENUATION)");
  return reinterpret_cast<realw*>(p);
}

static void ck_copy(Mesh* m, realw* dst, const realw* src, int64_t n) {
  if (n > 0 && dst && src)
    SF_CHECK(hipMemcpyAsync(dst, src, sizeof(realw) * (size_t)n, hipMemcpyDeviceToDevice, m->stream));
}

static void free_snapshot(Snapshot& s) {
  if (s.el) (void)hipFree(s.el);
  if (s.ac) (void)hipFree(s.ac);
  if (s.R) (void)hipFree(s.R);
  s = Snapshot();
}

void free_checkpoints(Mesh* m) {
  for (auto& s : m->snapshots) free_snapshot(s);
  m->snapshots.clear();
  if (m->history_el) (void)hipFree(m->history_el);
  if (m->history_ac) (void)hipFree(m->history_ac);
  m->history_el = m->history_ac = nullptr;
  m->history_frames = 0;
}

}  // namespace sf

using namespace sf;

// Save the state of one field (FORWARD_OR_ADJOINT 1 or 3) into snapshot slot `slot` (0-based,
// grown on demand). Like save_forward_arrays_undoatt: potentials, displ/veloc/accel and the
// memory variables; no strain.
extern "C" void sf_snapshot_save(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int slot) {
  Mesh* m = from_handle(Mesh_pointer);
  if (slot < 0) fatal("sf_snapshot_save", "negative slot");
  const int f = field_of_foa(FORWARD_OR_ADJOINT);
  flush_all(m);
  if ((int)m->snapshots.size() <= slot) m->snapshots.resize(slot + 1);
  Snapshot& s = m->snapshots[slot];
  const int64_t n = m->nglob, nloc = (int64_t)m->nspec * 125;
  if (m->el[f].displ) {
    if (!s.el) s.el = ck_alloc(9 * n, "sf_snapshot_save");
    ck_copy(m, s.el, m->el[f].displ, 3 * n);
    ck_copy(m, s.el + 3 * n, m->el[f].veloc, 3 * n);
    ck_copy(m, s.el + 6 * n, m->el[f].accel, 3 * n);
  }
  if (m->ac[f].pot) {
    if (!s.ac) s.ac = ck_alloc(3 * n, "sf_snapshot_save");
    ck_copy(m, s.ac, m->ac[f].pot, n);
    ck_copy(m, s.ac + n, m->ac[f].potd, n);
    ck_copy(m, s.ac + 2 * n, m->ac[f].potdd, n);
  }
  if (m->attenuation && m->att[f].R) {
    if (!s.R) s.R = ck_alloc(18 * nloc, "sf_snapshot_save");
    ck_copy(m, s.R, m->att[f].R, 18 * nloc);
  }
}

// Restore snapshot `slot` into one field (read_forward_arrays_undoatt). The field's strain is
// left as it was; the caller recomputes it (compute_strain_cuda) as the reference does.
extern "C" void sf_snapshot_load(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int slot) {
  Mesh* m = from_handle(Mesh_pointer);
  if (slot < 0 || slot >= (int)m->snapshots.size()) fatal("sf_snapshot_load", "snapshot slot was never saved");
  const int f = field_of_foa(FORWARD_OR_ADJOINT);
  flush_all(m);
  const Snapshot& s = m->snapshots[slot];
  const int64_t n = m->nglob, nloc = (int64_t)m->nspec * 125;
  if (m->el[f].displ) {
    if (!s.el) fatal("sf_snapshot_load", "snapshot holds no elastic field");
    ck_copy(m, m->el[f].displ, s.el, 3 * n);
    ck_copy(m, m->el[f].veloc, s.el + 3 * n, 3 * n);
    ck_copy(m, m->el[f].accel, s.el + 6 * n, 3 * n);
  }
  if (m->ac[f].pot) {
    if (!s.ac) fatal("sf_snapshot_load", "snapshot holds no acoustic field");
    ck_copy(m, m->ac[f].pot, s.ac, n);
    ck_copy(m, m->ac[f].potd, s.ac + n, n);
    ck_copy(m, m->ac[f].potdd, s.ac + 2 * n, n);
  }
  if (m->attenuation && m->att[f].R) {
    if (!s.R) fatal("sf_snapshot_load", "snapshot holds no memory variables");
    ck_copy(m, m->att[f].R, s.R, 18 * nloc);
  }
}

// Reserve `nframes` history frames (b_displ_elastic_buffer / b_potential_*_buffer of
// iterate_time_undoatt.F90:45-47); frees and reallocates when the size changes.
extern "C" void sf_history_reserve(long* Mesh_pointer, int nframes) {
  Mesh* m = from_handle(Mesh_pointer);
  if (nframes < 0) fatal("sf_history_reserve", "negative frame count");
  if (nframes == m->history_frames) return;
  SF_CHECK(hipStreamSynchronize(m->stream));
  if (m->history_el) (void)hipFree(m->history_el);
  if (m->history_ac) (void)hipFree(m->history_ac);
  m->history_el = m->history_ac = nullptr;
  m->history_frames = nframes;
  if (nframes == 0) return;
  const int64_t n = m->nglob;
  if (m->el[0].displ || m->el[1].displ) m->history_el = ck_alloc((int64_t)nframes * 3 * n, "sf_history_reserve");
  if (m->ac[0].pot || m->ac[1].pot) m->history_ac = ck_alloc((int64_t)nframes * 2 * n, "sf_history_reserve");
}

// frame <- displ (and potential, potential_dot_dot) of one field
extern "C" void sf_history_store(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int frame) {
  Mesh* m = from_handle(Mesh_pointer);
  if (frame < 0 || frame >= m->history_frames) fatal("sf_history_store", "frame outside the reserved history");
  const int f = field_of_foa(FORWARD_OR_ADJOINT);
  flush_all(m);
  const int64_t n = m->nglob;
  if (m->history_el && m->el[f].displ) ck_copy(m, m->history_el + (int64_t)frame * 3 * n, m->el[f].displ, 3 * n);
  if (m->history_ac && m->ac[f].pot) {
    ck_copy(m, m->history_ac + (int64_t)frame * 2 * n, m->ac[f].pot, n);
    ck_copy(m, m->history_ac + (int64_t)frame * 2 * n + n, m->ac[f].potdd, n);
  }
}

// displ (and potential, potential_dot_dot) of one field <- frame
extern "C" void sf_history_load(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int frame) {
  Mesh* m = from_handle(Mesh_pointer);
  if (frame < 0 || frame >= m->history_frames) fatal("sf_history_load", "frame outside the reserved history");
  const int f = field_of_foa(FORWARD_OR_ADJOINT);
  flush_all(m);
  const int64_t n = m->nglob;
  if (m->history_el && m->el[f].displ) ck_copy(m, m->el[f].displ, m->history_el + (int64_t)frame * 3 * n, 3 * n);
  if (m->history_ac && m->ac[f].pot) {
    ck_copy(m, m->ac[f].pot, m->history_ac + (int64_t)frame * 2 * n, n);
    ck_copy(m, m->ac[f].potdd, m->history_ac + (int64_t)frame * 2 * n + n, n);
  }
}

// zero the kernel accumulators of a kernel run (they are zeroed once by prepare_fields_*_adj_dev)
extern "C" void sf_zero_kernels(long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  const size_t b = sizeof(realw) * (size_t)m->nspec * 125;
  realw* k[5] = {m->rho_kl, m->mu_kl, m->kappa_kl, m->rho_ac_kl, m->kappa_ac_kl};
  for (realw* p : k)
    if (p) SF_CHECK(hipMemsetAsync(p, 0, b, m->stream));
}

// release every snapshot and the history (also done by prepare_cleanup_device)
extern "C" void sf_checkpoints_free(long* Mesh_pointer) {
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  free_checkpoints(m);
}
