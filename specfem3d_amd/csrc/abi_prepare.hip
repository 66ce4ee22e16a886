// abi_prepare.hip -- set-up / tear-down / transfer entry points of the Fortran-callable C-ABI.
//
// Signatures are those the reference's Fortran host binds (src/gpu/specfem3D_gpu_cuda_method_stubs.c
// :40-945; every argument by reference, Fortran logicals as int, arrays column-major with
// 1-based index values). The handle written to *Mesh_pointer is an sf::Mesh* owned by this
// library. Host arrays stay owned by the caller; everything is copied into device
// allocations laid out for gfx950 (see mesh.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace sf {
void free_checkpoints(Mesh* m);

Mesh* from_handle(long* p) {
  if (!p || !*p) fatal("Mesh_pointer", "null handle (prepare_constants_device not called?)");
  return reinterpret_cast<Mesh*>(*p);
}

// Set-up progress on stderr when SF_VERBOSE_PREPARE=1 (large meshes take minutes to prepare).
static void prep_log(const char* what) {
  static const bool on = [] {
    const char* v = std::getenv("SF_VERBOSE_PREPARE");
    return v && v[0] == '1';
  }();
  if (!on) return;
  static const auto t0 = std::chrono::steady_clock::now();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "[prepare] %s at %.1fs\n", what, s);
  std::fflush(stderr);
}

template <typename T>
static T* dmalloc(int64_t n) {
  if (n <= 0) n = 1;
  void* p = nullptr;
  SF_CHECK(hipMalloc(&p, sizeof(T) * (size_t)n));
  return reinterpret_cast<T*>(p);
}

template <typename T>
static T* dupload(const T* h, int64_t n) {
  T* d = dmalloc<T>(n);
  if (n > 0) SF_CHECK(hipMemcpy(d, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
  return d;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// host-side raw face data kept between prepare calls (faces need material arrays that arrive later)
struct RawFaces {
  std::vector<int> ispec;      // 0-based
  std::vector<int> ijk;        // 3*25*nf, 0-based
  std::vector<realw> normal;   // 3*25*nf
  std::vector<realw> jacw;     // 25*nf
};

struct HostSide {
  RawFaces abs;
  std::vector<int> ibool;      // 0-based
  std::vector<unsigned char> is_el, is_ac;
  std::vector<unsigned char> is_cpml;  // C-PML flag per element (prepare_fields_elastic_pml)
  int nspec = 0;
  // pinned staging ring for per-step host values (source time functions, adjoint sources)
  static const int RING = 64;
  realw* ring = nullptr;
  int ring_slot = 0;
  int ring_stride = 0;
  hipEvent_t ring_ev[RING];
  bool ring_ev_init = false;
  realw* stf_dev = nullptr;
  int nsrc_dev = 0;
  std::vector<int> src_ispec;  // element of each local source (0-based)
};

static HostSide* host_of(Mesh* m);

// Build face points from faces given as (ispec, ijk, normal, jacw) restricted to elements with
// flag[ispec] != 0, sorted into rounds (the r-th occurrence of a global point in face order).
static void build_face_points(Mesh* m, const RawFaces& rf, const unsigned char* flag, const int* ibool0,
                              FacePoints& fp, std::vector<int>* hp_ispec_out, std::vector<int>* hp_ploc_out) {
  const int nf = (int)rf.ispec.size();
  struct P { int g, e, p, slot, round, orig; realw n[3]; realw w; };
  std::vector<P> pts;
  pts.reserve((size_t)nf * 25);
  std::vector<int> occ;  // occurrence counter per global point (sparse via sort later)
  for (int f = 0; f < nf; f++) {
    const int e = rf.ispec[f];
    if (flag && !flag[e]) continue;
    for (int a = 0; a < 25; a++) {
      const int* ijk = &rf.ijk[((size_t)f * 25 + a) * 3];
      const int p = ijk[0] + 5 * ijk[1] + 25 * ijk[2];
      P q;
      q.g = ibool0[(int64_t)e * 125 + p];
      q.e = e;
      q.p = p;
      q.slot = f * 25 + a;
      q.round = 0;
      q.orig = (int)pts.size();
      for (int c = 0; c < 3; c++) q.n[c] = rf.normal.empty() ? 0.f : rf.normal[((size_t)f * 25 + a) * 3 + c];
      q.w = rf.jacw.empty() ? 0.f : rf.jacw[(size_t)f * 25 + a];
      pts.push_back(q);
    }
  }
  // round index = number of earlier occurrences of the same global point
  std::vector<std::pair<int, int>> key(pts.size());
  for (size_t i = 0; i < pts.size(); i++) key[i] = {pts[i].g, (int)i};
  std::sort(key.begin(), key.end());
  for (size_t i = 0; i < key.size(); i++) {
    if (i > 0 && key[i].first == key[i - 1].first)
      pts[key[i].second].round = pts[key[i - 1].second].round + 1;
  }
  int nr = 0;
  for (auto& q : pts) nr = std::max(nr, q.round + 1);
  if (nr > 16) fatal("build_face_points", "a global point belongs to more than 16 boundary faces");
  std::stable_sort(pts.begin(), pts.end(), [](const P& a, const P& b) { return a.round < b.round; });
  fp.n = (int)pts.size();
  fp.nrounds = nr;
  for (int r = 0; r <= 16; r++) fp.round_off[r] = 0;
  {
    int r = 0;
    for (int i = 0; i < fp.n; i++)
      while (r <= pts[i].round) fp.round_off[r++] = i;
    while (r <= nr) fp.round_off[r++] = fp.n;
  }
  // occurrence chains: the next occurrence of a point (key is in face order within each point)
  std::vector<int> pos(fp.n), nx(fp.n, -1);
  for (int i = 0; i < fp.n; i++) pos[pts[i].orig] = i;
  for (size_t i = 0; i + 1 < key.size(); i++)
    if (key[i + 1].first == key[i].first) nx[pos[key[i].second]] = pos[key[i + 1].second];
  std::vector<int> g(fp.n), e(fp.n), pl(fp.n), sl(fp.n);
  std::vector<realw> n3(3 * (size_t)fp.n), w(fp.n);
  for (int i = 0; i < fp.n; i++) {
    g[i] = pts[i].g; e[i] = pts[i].e; pl[i] = pts[i].p; sl[i] = pts[i].slot;
    for (int c = 0; c < 3; c++) n3[3 * (size_t)i + c] = pts[i].n[c];
    w[i] = pts[i].w;
  }
  fp.iglob = dupload(g.data(), fp.n);
  fp.ispec = dupload(e.data(), fp.n);
  fp.ploc = dupload(pl.data(), fp.n);
  fp.slot = dupload(sl.data(), fp.n);
  fp.normal = dupload(n3.data(), 3 * (int64_t)fp.n);
  fp.jacw = dupload(w.data(), fp.n);
  fp.next = dupload(nx.data(), fp.n);
  if (hp_ispec_out) *hp_ispec_out = e;
  if (hp_ploc_out) *hp_ploc_out = pl;
}

static void free_face_points(FacePoints& f) {
  dfree(f.iglob); dfree(f.ispec); dfree(f.ploc); dfree(f.slot); dfree(f.next);
  dfree(f.normal); dfree(f.jacw); dfree(f.imp_p); dfree(f.imp_s);
  f.n = 0; f.nrounds = 0;
}

// per-mesh host side storage (kept in a side table keyed by Mesh*)
static std::vector<std::pair<Mesh*, HostSide*>> g_host;
static HostSide* host_of(Mesh* m) {
  for (auto& p : g_host)
    if (p.first == m) return p.second;
  HostSide* h = new HostSide();
  g_host.push_back({m, h});
  return h;
}
static void drop_host(Mesh* m) {
  for (size_t i = 0; i < g_host.size(); i++)
    if (g_host[i].first == m) {
      HostSide* h = g_host[i].second;
      if (h->ring) (void)hipHostFree(h->ring);
      if (h->ring_ev_init)
        for (int r = 0; r < HostSide::RING; r++) (void)hipEventDestroy(h->ring_ev[r]);
      dfree(h->stf_dev);
      delete h;
      g_host.erase(g_host.begin() + i);
      return;
    }
}

// stage `n` floats from the host into a device buffer without racing earlier in-flight copies:
// a ring of pinned slots, each reused only after its previous copy has completed.
realw* stage_to_device(Mesh* m, const realw* vals, int n, realw* dst) {
  HostSide* h = host_of(m);
  if (n > h->ring_stride) {
    if (h->ring) {
      SF_CHECK(hipStreamSynchronize(m->stream));
      SF_CHECK(hipHostFree(h->ring));
    }
    h->ring_stride = std::max(n, 64);
    SF_CHECK(hipHostMalloc((void**)&h->ring, sizeof(realw) * (size_t)h->ring_stride * HostSide::RING,
                           hipHostMallocDefault));
    if (!h->ring_ev_init) {
      for (int r = 0; r < HostSide::RING; r++) SF_CHECK(hipEventCreateWithFlags(&h->ring_ev[r], hipEventDisableTiming));
      h->ring_ev_init = true;
    }
  }
  const int s = h->ring_slot;
  h->ring_slot = (s + 1) % HostSide::RING;
  SF_CHECK(hipEventSynchronize(h->ring_ev[s]));
  realw* slot = h->ring + (size_t)s * h->ring_stride;
  std::memcpy(slot, vals, sizeof(realw) * (size_t)n);
  SF_CHECK(hipMemcpyAsync(dst, slot, sizeof(realw) * (size_t)n, hipMemcpyHostToDevice, m->stream));
  SF_CHECK(hipEventRecord(h->ring_ev[s], m->stream));
  return dst;
}

realw* stf_device_buffer(Mesh* m, int n) {
  HostSide* h = host_of(m);
  if (n > h->nsrc_dev) {
    SF_CHECK(hipStreamSynchronize(m->stream));
    dfree(h->stf_dev);
    h->stf_dev = dmalloc<realw>(n);
    h->nsrc_dev = n;
  }
  return h->stf_dev;
}

const std::vector<int>& host_ibool(Mesh* m) { return host_of(m)->ibool; }

// (owner, point) pairs in owner order for sources / adjoint sources, with rounds
static void build_point_sets(const std::vector<int>& owner_ispec, const std::vector<int>& ibool0, FacePoints& fp) {
  RawFaces rf;  // reuse the face builder: 5 "faces" of 25 points per owner element
  const int n = (int)owner_ispec.size();
  for (int s = 0; s < n; s++) {
    for (int f5 = 0; f5 < 5; f5++) {
      rf.ispec.push_back(owner_ispec[s]);
      for (int a = 0; a < 25; a++) {
        const int p = f5 * 25 + a;
        rf.ijk.push_back(p % 5);
        rf.ijk.push_back((p / 5) % 5);
        rf.ijk.push_back(p / 25);
      }
    }
  }
  std::vector<int> es;
  build_face_points(nullptr, rf, nullptr, ibool0.data(), fp, &es, nullptr);
  // owner index instead of element: slot / 125 (5 faces x 25 points per owner)
  std::vector<int> slots(fp.n), own(fp.n);
  SF_CHECK(hipMemcpy(slots.data(), fp.slot, sizeof(int) * (size_t)fp.n, hipMemcpyDeviceToHost));
  for (int i = 0; i < fp.n; i++) own[i] = slots[i] / 125;
  SF_CHECK(hipMemcpy(fp.ispec, own.data(), sizeof(int) * (size_t)fp.n, hipMemcpyHostToDevice));
}

}  // namespace sf

using namespace sf;

// --------------------------------------------------------------------------------------------
// initialize_gpu.cu:535
extern "C" void FC_FUNC_(initialize_gpu_device, INITIALIZE_GPU_DEVICE)(int* myrank_f, int* ncuda_devices) {
  int n = 0;
  SF_CHECK(hipGetDeviceCount(&n));
  if (n <= 0) fatal("initialize_gpu_device", "no HIP device found");
  const int dev = (myrank_f ? *myrank_f : 0) % n;
  SF_CHECK(hipSetDevice(dev));
  if (ncuda_devices) *ncuda_devices = n;
}

// prepare_mesh_constants_cuda.cu:70
extern "C" void FC_FUNC_(prepare_constants_device, PREPARE_CONSTANTS_DEVICE)(
    long* Mesh_pointer, int* h_NGLLX, int* NSPEC_AB, int* NGLOB_AB, int* NSPEC_IRREGULAR,
    int* h_irregular_element_number, realw* h_xix, realw* h_xiy, realw* h_xiz, realw* h_etax, realw* h_etay,
    realw* h_etaz, realw* h_gammax, realw* h_gammay, realw* h_gammaz, realw* xix_regular, realw* jacobian_regular,
    int* h_ibool, int* num_interfaces_ext_mesh, int* max_nibool_interfaces_ext_mesh,
    int* h_nibool_interfaces_ext_mesh, int* h_ibool_interfaces_ext_mesh, realw* h_hprime_xx,
    realw* h_hprimewgll_xx, realw* h_wgllwgll_xy, realw* h_wgllwgll_xz, realw* h_wgllwgll_yz,
    int* STACEY_ABSORBING_CONDITIONS, int* h_abs_boundary_ispec, int* h_abs_boundary_ijk,
    realw* h_abs_boundary_normal, realw* h_abs_boundary_jacobian2Dw, int* h_num_abs_boundary_faces,
    int* h_ispec_is_inner, int* NSOURCES, int* nsources_local_f, realw* h_sourcearrays,
    int* h_islice_selected_source, int* h_ispec_selected_source, int* h_ispec_selected_rec, int* nrec,
    int* nrec_local, int* SIMULATION_TYPE, int* USE_MESH_COLORING_GPU_f, int* nspec_acoustic, int* nspec_elastic,
    int* ispec_is_acoustic, int* ispec_is_elastic, int* h_myrank, int* SAVE_FORWARD, realw* h_xir, realw* h_etar,
    realw* h_gammar, double* nu_rec, double* nu_source, int* h_islice_selected_rec, int* nlength_seismogram,
    int* SAVE_SEISMOGRAMS_DISPLACEMENT, int* SAVE_SEISMOGRAMS_VELOCITY, int* SAVE_SEISMOGRAMS_ACCELERATION,
    int* SAVE_SEISMOGRAMS_PRESSURE, int* h_NB_RUNS_ACOUSTIC_GPU, int* FAULT_SIMULATION,
    int* UNDO_ATTENUATION_AND_OR_PML, int* PML_CONDITIONS) {
  if (*h_NGLLX != NGLLX) fatal("prepare_constants_device", "this library is built for NGLLX = 5");
  if (h_NB_RUNS_ACOUSTIC_GPU && *h_NB_RUNS_ACOUSTIC_GPU != 1)
    fatal("prepare_constants_device", "NB_RUNS_ACOUSTIC_GPU > 1 is not supported (use the batched engine)");
  if (FAULT_SIMULATION && *FAULT_SIMULATION) fatal("prepare_constants_device", "fault simulations are not supported");
  if (num_interfaces_ext_mesh && *num_interfaces_ext_mesh != 0)
    fatal("prepare_constants_device", "MPI interfaces present: this library runs one partition (NPROC=1)");
  (void)max_nibool_interfaces_ext_mesh; (void)h_nibool_interfaces_ext_mesh; (void)h_ibool_interfaces_ext_mesh;
  (void)h_ispec_is_inner; (void)USE_MESH_COLORING_GPU_f; (void)nu_source; (void)h_wgllwgll_xz; (void)h_wgllwgll_yz;

  Mesh* m = new Mesh();
  *Mesh_pointer = reinterpret_cast<long>(m);
  HostSide* H = host_of(m);
  SF_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  m->nspec = *NSPEC_AB;
  m->nglob = *NGLOB_AB;
  m->nspec_irregular = *NSPEC_IRREGULAR;
  m->simulation_type = *SIMULATION_TYPE;
  m->save_forward = *SAVE_FORWARD;
  m->myrank = *h_myrank;
  m->xix_regular = *xix_regular;
  m->jacobian_regular = *jacobian_regular;
  H->nspec = m->nspec;
  if (const char* e = std::getenv("SPECFEM_AMD_EXACT")) m->exact = std::atoi(e) != 0;
  if (const char* e = std::getenv("SPECFEM_AMD_ASSEMBLY")) m->assembly = std::atoi(e);

  // GLL constants: hp[i*5+l] = hprime_xx(i+1,l+1) (column-major offset i + 5 l)
  realw hp[25], hpw[25], wxy[25];
  for (int i = 0; i < 5; i++)
    for (int l = 0; l < 5; l++) {
      hp[i * 5 + l] = h_hprime_xx[i + 5 * l];
      hpw[i * 5 + l] = h_hprimewgll_xx[i + 5 * l];
      wxy[i * 5 + l] = h_wgllwgll_xy[i + 5 * l];
    }
  set_gll_constants(hp, hpw, wxy);

  const int64_t nloc = (int64_t)m->nspec * 125;
  H->ibool.resize(nloc);
  for (int64_t q = 0; q < nloc; q++) H->ibool[q] = h_ibool[q] - 1;
  m->ibool = dupload(H->ibool.data(), nloc);
  prep_log("ibool uploaded");
  m->irregular = dupload(h_irregular_element_number, m->nspec);

  // irregular Jacobians: 9 planes from the host + the determinant recomputed as the inverse of
  // det(d xi / d x) (the host does not pass jacobianstore through this interface)
  const int64_t P = (int64_t)m->nspec_irregular * 125;
  if (P > 0) {
    std::vector<realw> J(10 * (size_t)P);
    const realw* src[9] = {h_xix, h_xiy, h_xiz, h_etax, h_etay, h_etaz, h_gammax, h_gammay, h_gammaz};
    for (int t = 0; t < 9; t++) std::memcpy(&J[(size_t)t * P], src[t], sizeof(realw) * (size_t)P);
    for (int64_t q = 0; q < P; q++) {
      const realw xixl = J[q], xiyl = J[P + q], xizl = J[2 * P + q], etaxl = J[3 * P + q], etayl = J[4 * P + q];
      const realw etazl = J[5 * P + q], gaxl = J[6 * P + q], gayl = J[7 * P + q], gazl = J[8 * P + q];
      J[9 * P + q] = 1.f / (xixl * (etayl * gazl - etazl * gayl) - xiyl * (etaxl * gazl - etazl * gaxl) +
                            xizl * (etaxl * gayl - etayl * gaxl));
    }
    m->jac = dupload(J.data(), 10 * P);
  } else {
    m->jac = dmalloc<realw>(10);
  }

  // domains
  H->is_el.resize(m->nspec);
  H->is_ac.resize(m->nspec);
  for (int e = 0; e < m->nspec; e++) {
    H->is_el[e] = ispec_is_elastic[e] != 0;
    H->is_ac[e] = ispec_is_acoustic[e] != 0;
  }
  m->nspec_elastic = *nspec_elastic;
  m->nspec_acoustic = *nspec_acoustic;

  // absorbing faces (kept on the host until the material arrays arrive); with C-PML they are the
  // outer faces of the layers, which carry its Dirichlet condition
  m->stacey = *STACEY_ABSORBING_CONDITIONS;
  m->pml_conditions = PML_CONDITIONS ? *PML_CONDITIONS : 0;
  if (m->pml_conditions && m->stacey) fatal("prepare_constants_device", "C-PML and Stacey conditions exclude each other");
  m->num_abs_boundary_faces = *h_num_abs_boundary_faces;
  if (m->stacey || m->pml_conditions) {
    const int nf = m->num_abs_boundary_faces;
    H->abs.ispec.resize(nf);
    H->abs.ijk.resize((size_t)nf * 75);
    H->abs.normal.assign(h_abs_boundary_normal, h_abs_boundary_normal + (size_t)nf * 75);
    H->abs.jacw.assign(h_abs_boundary_jacobian2Dw, h_abs_boundary_jacobian2Dw + (size_t)nf * 25);
    for (int f = 0; f < nf; f++) H->abs.ispec[f] = h_abs_boundary_ispec[f] - 1;
    for (size_t q = 0; q < (size_t)nf * 75; q++) H->abs.ijk[q] = h_abs_boundary_ijk[q] - 1;
  }

  // point sources in this slice: sourcearrays(NSOURCES, 3, 5, 5, 5) column-major
  const int ns = *NSOURCES;
  std::vector<int> src_ispec;
  std::vector<realw> arr;
  for (int s = 0; s < ns; s++) {
    if (h_islice_selected_source[s] != m->myrank) continue;
    const int e = h_ispec_selected_source[s] - 1;
    src_ispec.push_back(e);
    for (int p = 0; p < 125; p++)
      for (int c = 0; c < 3; c++) arr.push_back(h_sourcearrays[s + (int64_t)ns * (c + 3 * p)]);
  }
  m->nsources = (int)src_ispec.size();
  H->src_ispec = src_ispec;

  (void)nsources_local_f;
  // sources are added in the domain of their element: elastic sources to accel
  // (compute_add_sources_viscoelastic.F90:143), acoustic ones to potential_dot_dot_acoustic
  // (compute_add_sources_acoustic.f90:97)
  {
    m->src_array = dupload(arr.data(), (int64_t)arr.size());
    m->h_src_elastic = new unsigned char[std::max(1, m->nsources)];
    m->h_src_acoustic = new unsigned char[std::max(1, m->nsources)];
    std::vector<int> own_el, idx_el, own_ac, idx_ac;
    for (int s = 0; s < m->nsources; s++) {
      m->h_src_elastic[s] = H->is_el[src_ispec[s]];
      m->h_src_acoustic[s] = H->is_ac[src_ispec[s]];
      if (m->h_src_elastic[s]) {
        own_el.push_back(src_ispec[s]);
        idx_el.push_back(s);
      }
      if (m->h_src_acoustic[s]) {
        own_ac.push_back(src_ispec[s]);
        idx_ac.push_back(s);
      }
    }
    // owner index = position in the domain's source list -> index into the full source list
    auto remap_owner = [](FacePoints& fp, const std::vector<int>& idx) {
      std::vector<int> own(fp.n);
      SF_CHECK(hipMemcpy(own.data(), fp.ispec, sizeof(int) * own.size(), hipMemcpyDeviceToHost));
      for (auto& o : own) o = idx[o];
      SF_CHECK(hipMemcpy(fp.ispec, own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice));
    };
    if (!own_el.empty()) {
      build_point_sets(own_el, H->ibool, m->src_pts);
      remap_owner(m->src_pts, idx_el);
    }
    if (!own_ac.empty()) {
      build_point_sets(own_ac, H->ibool, m->src_ac_pts);
      remap_owner(m->src_ac_pts, idx_ac);
    }
  }

  // receivers local to this slice: hxir_store(5, nrec_local) etc. are CUSTOM_REAL
  m->nrec_local = *nrec_local;
  m->nlength_seismogram = *nlength_seismogram;
  m->save_d = *SAVE_SEISMOGRAMS_DISPLACEMENT;
  m->save_v = *SAVE_SEISMOGRAMS_VELOCITY;
  m->save_a = *SAVE_SEISMOGRAMS_ACCELERATION;
  m->save_p = *SAVE_SEISMOGRAMS_PRESSURE;
  if (m->nrec_local > 0) {
    std::vector<int> ig((size_t)m->nrec_local * 125);
    std::vector<double> h((size_t)m->nrec_local * 125), nu((size_t)m->nrec_local * 9);
    m->h_rec_ispec = new int[m->nrec_local];
    m->h_rec_elastic = new unsigned char[m->nrec_local];
    std::vector<int> el_list, ac_list;
    int irl = 0;
    for (int r = 0; r < *nrec && irl < m->nrec_local; r++) {
      if (h_islice_selected_rec[r] != m->myrank) continue;
      const int e = h_ispec_selected_rec[r] - 1;
      m->h_rec_ispec[irl] = e;
      m->h_rec_elastic[irl] = H->is_el[e];
      if (H->is_el[e]) el_list.push_back(irl);
      else if (H->is_ac[e]) ac_list.push_back(irl);
      for (int p = 0; p < 125; p++) {
        const int i = p % 5, j = (p / 5) % 5, k = p / 25;
        ig[(size_t)irl * 125 + p] = H->ibool[(int64_t)e * 125 + p];
        h[(size_t)irl * 125 + p] = (double)h_xir[i + 5 * irl] * (double)h_etar[j + 5 * irl] * (double)h_gammar[k + 5 * irl];
      }
      for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) nu[(size_t)irl * 9 + 3 * c + d] = nu_rec[c + 3 * d + 9 * r];
      irl++;
    }
    m->rec_iglob = dupload(ig.data(), (int64_t)ig.size());
    m->rec_ispec = dupload(m->h_rec_ispec, (int64_t)m->nrec_local);
    m->nrec_el = (int)el_list.size();
    m->nrec_ac = (int)ac_list.size();
    m->rec_el_list = dupload(el_list.data(), (int64_t)el_list.size());
    m->rec_ac_list = dupload(ac_list.data(), (int64_t)ac_list.size());
    m->rec_h = dupload(h.data(), (int64_t)h.size());
    m->rec_nu = dupload(nu.data(), (int64_t)nu.size());
    const int64_t sz = (int64_t)m->nlength_seismogram * m->nrec_local * 3;
    if (m->save_d) m->seismo_d = dmalloc<realw>(sz);
    if (m->save_v) m->seismo_v = dmalloc<realw>(sz);
    if (m->save_a) m->seismo_a = dmalloc<realw>(sz);
    if (m->save_p) m->seismo_p = dmalloc<realw>(sz);
  }
  m->undo_attenuation = *UNDO_ATTENUATION_AND_OR_PML;
  m->d_norm = dmalloc<realw>(4);
  SF_CHECK(hipHostMalloc((void**)&m->h_norm, sizeof(realw) * 4, hipHostMallocDefault));
}

namespace sf {
static void build_elastic_assembly(Mesh* m);

// Distinct (factor_common, factor_common_kappa) rows, compared bit for bit, and a 16-bit row index
// per local point; gives up (planes only) beyond 65535 distinct rows.
static void build_factor_table(Mesh* m, const realw* fc, const realw* fck, int64_t nloc) {
  struct Key {
    uint32_t w[6];
    bool operator==(const Key& o) const { return std::memcmp(w, o.w, sizeof(w)) == 0; }
  };
  struct Hash {
    size_t operator()(const Key& k) const {
      uint64_t h = 1469598103934665603ull;
      for (uint32_t v : k.w) h = (h ^ v) * 1099511628211ull;
      return (size_t)h;
    }
  };
  std::unordered_map<Key, int, Hash> rows;
  std::vector<Key> order;
  std::vector<uint16_t> idx((size_t)nloc);
  Key last{};
  int last_id = -1;
  for (int64_t q = 0; q < nloc; q++) {
    Key k;
    std::memcpy(k.w, fc + 3 * q, 12);
    std::memcpy(k.w + 3, fck + 3 * q, 12);
    if (last_id >= 0 && k == last) {
      idx[q] = (uint16_t)last_id;
      continue;
    }
    auto it = rows.find(k);
    int id;
    if (it == rows.end()) {
      if (order.size() >= 65535) return;  // too many distinct rows: keep the planes
      id = (int)order.size();
      rows.emplace(k, id);
      order.push_back(k);
    } else {
      id = it->second;
    }
    idx[q] = (uint16_t)id;
    last = k;
    last_id = id;
  }
  const int n = (int)order.size();
  std::vector<realw> tab(6 * (size_t)n);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < 6; c++) std::memcpy(&tab[(size_t)c * n + r], &order[r].w[c], 4);
  m->fc_table = dupload(tab.data(), 6 * (int64_t)n);
  m->fc_index = dupload(idx.data(), nloc);
  m->fc_rows = n;
}

static void build_colors(Mesh* m, const std::vector<int>& elems, int*& d_list, int* off, int& ncolor);
}  // namespace sf

namespace sf {
// Host arrays in the reference's layouts -> device per-element blocks (r_idx / eps_idx):
// epsilondev_*(NGLL3, NSPEC), one array per component (a null component stays zero), and
// R_*(N_SLS, NGLL3, NSPEC), SLS fastest.
static realw* upload_strain_blocks(const realw* const src[6], int64_t nloc) {
  std::vector<realw> E(6 * (size_t)nloc, 0.f);
  const int64_t ne = nloc / 125;
  for (int c = 0; c < 6; c++) {
    if (!src[c]) continue;
    for (int64_t e = 0; e < ne; e++)
      std::memcpy(&E[(size_t)eps_idx(e, c, 0)], src[c] + e * 125, sizeof(realw) * 125);
  }
  return dupload(E.data(), 6 * nloc);
}
static realw* upload_memory_blocks(const realw* const rs[6], int64_t nloc) {
  std::vector<realw> R(18 * (size_t)nloc);
  const int64_t ne = nloc / 125;
  for (int64_t e = 0; e < ne; e++)
    for (int c = 0; c < 6; c++) {
      const realw* src = rs[c] + 3 * e * 125;
      for (int s = 0; s < 3; s++) {
        realw* dst = &R[(size_t)r_idx(e, c * 3 + s, 0)];
        for (int p = 0; p < 125; p++) dst[p] = src[3 * p + s];
      }
    }
  return dupload(R.data(), 18 * nloc);
}
}  // namespace sf

// prepare_mesh_constants_cuda.cu:657
extern "C" void FC_FUNC_(prepare_fields_elastic_device, PREPARE_FIELDS_ELASTIC_DEVICE)(
    long* Mesh_pointer, realw* rmassx, realw* rmassy, realw* rmassz, realw* rho_vp, realw* rho_vs, realw* h_kappav,
    realw* h_muv, int* num_phase_ispec_elastic, int* phase_ispec_inner_elastic, realw* b_absorb_field,
    int* b_reclen_field, int* COMPUTE_AND_STORE_STRAIN, realw* epsilondev_xx, realw* epsilondev_yy,
    realw* epsilondev_xy, realw* epsilondev_xz, realw* epsilondev_yz, int* ATTENUATION, int* R_size, realw* R_xx,
    realw* R_yy, realw* R_xy, realw* R_xz, realw* R_yz, realw* factor_common, realw* R_trace,
    realw* epsilondev_trace, realw* factor_common_kappa, realw* alphaval, realw* betaval, realw* gammaval,
    int* APPROXIMATE_OCEAN_LOAD, realw* rmass_ocean_load, int* NOISE_TOMOGRAPHY, realw* free_surface_normal,
    int* free_surface_ispec, int* free_surface_ijk, int* num_free_surface_faces, int* ACOUSTIC_SIMULATION,
    int* num_colors_outer_elastic, int* num_colors_inner_elastic, int* num_elem_colors_elastic, int* ANISOTROPY,
    realw* c11store, realw* c12store, realw* c13store, realw* c14store, realw* c15store, realw* c16store,
    realw* c22store, realw* c23store, realw* c24store, realw* c25store, realw* c26store, realw* c33store,
    realw* c34store, realw* c35store, realw* c36store, realw* c44store, realw* c45store, realw* c46store,
    realw* c55store, realw* c56store, realw* c66store) {
  Mesh* m = from_handle(Mesh_pointer);
  HostSide* H = host_of(m);
  if (*APPROXIMATE_OCEAN_LOAD) fatal("prepare_fields_elastic_device", "ocean load is not supported");
  if (*ANISOTROPY) fatal("prepare_fields_elastic_device", "anisotropy is not supported");
  if (*NOISE_TOMOGRAPHY) fatal("prepare_fields_elastic_device", "noise tomography is not supported");
  (void)rmass_ocean_load; (void)free_surface_normal; (void)free_surface_ispec; (void)free_surface_ijk;
  (void)num_free_surface_faces; (void)ACOUSTIC_SIMULATION; (void)num_colors_outer_elastic;
  (void)num_colors_inner_elastic; (void)num_elem_colors_elastic; (void)b_reclen_field;
  (void)c11store; (void)c12store; (void)c13store; (void)c14store; (void)c15store; (void)c16store; (void)c22store;
  (void)c23store; (void)c24store; (void)c25store; (void)c26store; (void)c33store; (void)c34store; (void)c35store;
  (void)c36store; (void)c44store; (void)c45store; (void)c46store; (void)c55store; (void)c56store; (void)c66store;

  const int64_t n = m->nglob;
  const int64_t nloc = (int64_t)m->nspec * 125;
  if (m->assembly < 0) {
    // automatic assembly: coloured without attenuation, atomic with it (see Mesh::assembly)
    m->assembly = *ATTENUATION ? 0 : 2;
    if (m->assembly == 2 && m->nspec_acoustic > 0 && m->acoustic_elems && !m->ac_color_elems) {
      std::vector<int> lst(m->nspec_acoustic);
      SF_CHECK(hipMemcpy(lst.data(), m->acoustic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
      build_colors(m, lst, m->ac_color_elems, m->ac_color_off, m->ac_ncolor);
    }
  }
  {
    std::vector<realw> rm(3 * (size_t)n);
    for (int64_t g = 0; g < n; g++) {
      rm[3 * g] = rmassx[g];
      rm[3 * g + 1] = rmassy[g];
      rm[3 * g + 2] = rmassz[g];
    }
    m->rmass = dupload(rm.data(), 3 * n);
  }
  prep_log("elastic mass uploaded");
  {
    std::vector<float2> km(nloc);
    for (int64_t q = 0; q < nloc; q++) km[q] = make_float2(h_kappav[q], h_muv[q]);
    m->kappamu = dupload(km.data(), nloc);
  }
  prep_log("elastic moduli uploaded");
  // element list: phase_ispec_inner_elastic(num_phase, 2); at NPROC=1 column 2 holds all elements
  {
    const int np = *num_phase_ispec_elastic;
    std::vector<int> lst;
    for (int ph = 0; ph < 2; ph++)
      for (int q = 0; q < np; q++) {
        const int e = phase_ispec_inner_elastic[q + (int64_t)np * ph];
        if (e > 0) lst.push_back(e - 1);
      }
    m->nspec_elastic = (int)lst.size();
    m->elastic_elems = dupload(lst.data(), (int64_t)lst.size());
    int gmin = m->nglob, gmax = -1;
    for (int e : lst)
      for (int p = 0; p < 125; p++) {
        const int g = H->ibool[(int64_t)e * 125 + p];
        gmin = std::min(gmin, g);
        gmax = std::max(gmax, g);
      }
    m->el_g0 = gmax < 0 ? 0 : (gmin / 4) * 4;
    m->el_n = gmax < 0 ? 0 : (int64_t)gmax + 1 - m->el_g0;
  }
  m->el[0].displ = dmalloc<realw>(3 * n);
  m->el[0].veloc = dmalloc<realw>(3 * n);
  m->el[0].accel = dmalloc<realw>(3 * n);
  SF_CHECK(hipMemset(m->el[0].displ, 0, sizeof(realw) * 3 * n));
  SF_CHECK(hipMemset(m->el[0].veloc, 0, sizeof(realw) * 3 * n));
  SF_CHECK(hipMemset(m->el[0].accel, 0, sizeof(realw) * 3 * n));

  // Stacey face points in elastic elements with rho*vp, rho*vs gathered per point
  if (m->stacey && m->num_abs_boundary_faces > 0) {
    std::vector<int> es, pl;
    build_face_points(m, H->abs, H->is_el.data(), H->ibool.data(), m->abs_el, &es, &pl);
    prep_log("elastic Stacey points built");
    std::vector<realw> a(m->abs_el.n), b(m->abs_el.n);
    for (int i = 0; i < m->abs_el.n; i++) {
      a[i] = rho_vp[(int64_t)es[i] * 125 + pl[i]];
      b[i] = rho_vs[(int64_t)es[i] * 125 + pl[i]];
    }
    m->abs_el.imp_p = dupload(a.data(), m->abs_el.n);
    m->abs_el.imp_s = dupload(b.data(), m->abs_el.n);
    if (m->save_forward || m->simulation_type == 3)
      m->b_absorb_field = dmalloc<realw>(3 * 25 * (int64_t)std::max(1, m->num_abs_boundary_faces));
  }
  (void)b_absorb_field;

  m->compute_and_store_strain = *COMPUTE_AND_STORE_STRAIN;
  m->attenuation = *ATTENUATION;
  if (m->compute_and_store_strain || m->attenuation) {
    const realw* src[6] = {epsilondev_xx, epsilondev_yy, epsilondev_xy, epsilondev_xz, epsilondev_yz,
                           m->attenuation ? epsilondev_trace : nullptr};
    m->att[0].eps = upload_strain_blocks(src, nloc);
    prep_log("strain uploaded");
  }
  if (m->attenuation) {
    if (*R_size != 3 * nloc) fatal("prepare_fields_elastic_device", "R_size must be N_SLS*125*NSPEC");
    const realw* rs[6] = {R_xx, R_yy, R_xy, R_xz, R_yz, R_trace};
    m->att[0].R = upload_memory_blocks(rs, nloc);
    prep_log("memory variables uploaded");
    std::vector<realw> F(3 * (size_t)nloc), Fk(3 * (size_t)nloc);
    for (int64_t q = 0; q < nloc; q++)
      for (int s = 0; s < 3; s++) {
        F[(size_t)s * nloc + q] = factor_common[3 * q + s];
        Fk[(size_t)s * nloc + q] = factor_common_kappa[3 * q + s];
      }
    m->factor_common = dupload(F.data(), 3 * nloc);
    m->factor_common_kappa = dupload(Fk.data(), 3 * nloc);
    prep_log("attenuation factors uploaded");
    build_factor_table(m, factor_common, factor_common_kappa, nloc);
    prep_log("attenuation factor table built");
    for (int s = 0; s < 3; s++) {
      m->alphaval[s] = alphaval[s];
      m->betaval[s] = betaval[s];
      m->gammaval[s] = gammaval[s];
    }
  }
  if (m->assembly == 1) build_elastic_assembly(m);
  if (m->assembly == 2) {
    std::vector<int> lst(m->nspec_elastic);
    SF_CHECK(hipMemcpy(lst.data(), m->elastic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
    build_colors(m, lst, m->el_color_elems, m->el_color_off, m->el_ncolor);
  }
}

namespace sf {
// elastic element colours for coloured launches that are not the mesh's own assembly mode
// (the batched engine)
void ensure_element_colors(Mesh* m) {
  if (m->el_color_elems || m->nspec_elastic <= 0 || !m->elastic_elems) return;
  std::vector<int> lst(m->nspec_elastic);
  const int* src = m->elastic_launch ? m->elastic_launch : m->elastic_elems;  // same colours as the mesh's own
  SF_CHECK(hipMemcpy(lst.data(), src, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
  build_colors(m, lst, m->el_color_elems, m->el_color_off, m->el_ncolor);
}

// acoustic element colours for the batched engine (the same colours a coloured single-shot run
// builds: greedy first-fit over the acoustic element list in its processing order)
void ensure_acoustic_colors(Mesh* m) {
  if (m->ac_color_elems || m->nspec_acoustic <= 0 || !m->acoustic_elems) return;
  std::vector<int> lst(m->nspec_acoustic);
  SF_CHECK(hipMemcpy(lst.data(), m->acoustic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
  build_colors(m, lst, m->ac_color_elems, m->ac_color_off, m->ac_ncolor);
}

// domain of element e (0-based) as prepare_constants_device received it
bool element_is_elastic(Mesh* m, int e) {
  const HostSide* H = host_of(m);
  return e >= 0 && e < (int)H->is_el.size() && H->is_el[e];
}

// Greedy first-fit colouring of an element list (in list order): an element takes the lowest
// colour not yet used by any element sharing one of its global points. Each colour is then one
// launch whose element-to-global scatter needs no atomics. Elements keep their list order
// inside a colour. At most 32 colours (a bit mask per global point); a structured hexahedral
// mesh needs 8.
static int color_elements(int nglob, int nelem, const int* ibool0, const int* elems, int* col) {
  std::vector<uint32_t> used((size_t)nglob, 0u);
  int ncolor = 0;
  for (int q = 0; q < nelem; q++) {
    const int* ib = ibool0 + (int64_t)elems[q] * 125;
    uint32_t u = 0;
    for (int p = 0; p < 125; p++) u |= used[ib[p]];
    if (u == 0xffffffffu) return -1;
    const int c = __builtin_ctz(~u);
    col[q] = c;
    for (int p = 0; p < 125; p++) used[ib[p]] |= 1u << c;
    if (c + 1 > ncolor) ncolor = c + 1;
  }
  return ncolor;
}

static void build_colors(Mesh* m, const std::vector<int>& elems, int*& d_list, int* off, int& ncolor) {
  HostSide* H = host_of(m);
  std::vector<int> col(elems.size());
  ncolor = color_elements(m->nglob, (int)elems.size(), H->ibool.data(), elems.data(), col.data());
  if (ncolor < 0) fatal("build_colors", "element colouring needs more than 32 colours");
  m->color_version++;
  for (int c = 0; c <= 32; c++) off[c] = 0;
  for (size_t q = 0; q < elems.size(); q++) off[col[q] + 1]++;
  for (int c = 0; c < 32; c++) off[c + 1] += off[c];
  std::vector<int> pos(off, off + 32), lst(elems.size());
  for (size_t q = 0; q < elems.size(); q++) lst[pos[col[q]]++] = elems[q];
  if (d_list) dfree(d_list);
  d_list = lst.empty() ? nullptr : dupload(lst.data(), (int64_t)lst.size());
}

// CSR of local slots per global point sorted by the position of their element in the
// processing list (ordered assembly); shared by the elastic and acoustic gathers.
static void build_csr(Mesh* m, const std::vector<int>& elems, int*& ptr, int*& idx) {
  HostSide* H = host_of(m);
  const int n = m->nglob;
  std::vector<int> cnt(n + 1, 0);
  for (int e : elems)
    for (int p = 0; p < 125; p++) cnt[H->ibool[(int64_t)e * 125 + p] + 1]++;
  for (int g = 0; g < n; g++) cnt[g + 1] += cnt[g];
  std::vector<int> id(cnt[n]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int e : elems)
    for (int p = 0; p < 125; p++) {
      const int g = H->ibool[(int64_t)e * 125 + p];
      id[pos[g]++] = e * 125 + p;
    }
  ptr = dupload(cnt.data(), n + 1);
  idx = dupload(id.data(), (int64_t)id.size());
}

static void build_elastic_assembly(Mesh* m) {
  // the stiffness elements, then the C-PML elements (compute_forces_viscoelastic, then
  // compute_forces_viscoelastic_PML: that is the order their contributions are summed in)
  std::vector<int> lst(m->nspec_elastic + m->pml_n);
  SF_CHECK(hipMemcpy(lst.data(), m->elastic_elems, sizeof(int) * m->nspec_elastic, hipMemcpyDeviceToHost));
  if (m->pml_n)
    SF_CHECK(hipMemcpy(lst.data() + m->nspec_elastic, m->pml_elems, sizeof(int) * m->pml_n, hipMemcpyDeviceToHost));
  dfree(m->asm_ptr);
  dfree(m->asm_idx);
  build_csr(m, lst, m->asm_ptr, m->asm_idx);
  if (!m->asm_buf) m->asm_buf = dmalloc<realw>(3 * (int64_t)m->nspec * 125);
  m->asm_ready = 1;
}


}  // namespace sf

// prepare_mesh_constants_cuda.cu:434
extern "C" void FC_FUNC_(prepare_fields_acoustic_device, PREPARE_FIELDS_ACOUSTIC_DEVICE)(
    long* Mesh_pointer, realw* rmass_acoustic, realw* rhostore, realw* kappastore, int* num_phase_ispec_acoustic,
    int* phase_ispec_inner_acoustic, int* NOISE_TOMOGRAPHY, int* num_free_surface_faces, int* free_surface_ispec,
    int* free_surface_ijk, int* b_reclen_potential, realw* b_absorb_potential, int* ELASTIC_SIMULATION,
    int* num_coupling_ac_el_faces, int* coupling_ac_el_ispec, int* coupling_ac_el_ijk, realw* coupling_ac_el_normal,
    realw* coupling_ac_el_jacobian2Dw, int* num_colors_outer_acoustic, int* num_colors_inner_acoustic,
    int* num_elem_colors_acoustic) {
  Mesh* m = from_handle(Mesh_pointer);
  HostSide* H = host_of(m);
  if (*NOISE_TOMOGRAPHY) fatal("prepare_fields_acoustic_device", "noise tomography is not supported");
  (void)b_reclen_potential; (void)b_absorb_potential; (void)ELASTIC_SIMULATION; (void)num_colors_outer_acoustic;
  (void)num_colors_inner_acoustic; (void)num_elem_colors_acoustic;
  const int64_t n = m->nglob, nloc = (int64_t)m->nspec * 125;
  m->rmass_ac = dupload(rmass_acoustic, n);
  m->rho = dupload(rhostore, nloc);
  m->kappa_ac = dupload(kappastore, nloc);
  if (m->src_ac_pts.n > 0) {
    // kappastore at the points of each source element (the source term is divided by it)
    std::vector<realw> k((size_t)m->nsources * 125);
    for (int src = 0; src < m->nsources; src++)
      for (int p = 0; p < 125; p++) k[(size_t)src * 125 + p] = kappastore[(int64_t)H->src_ispec[src] * 125 + p];
    m->src_kappa = dupload(k.data(), (int64_t)k.size());
  }
  {
    const int np = *num_phase_ispec_acoustic;
    std::vector<int> lst;
    for (int ph = 0; ph < 2; ph++)
      for (int q = 0; q < np; q++) {
        const int e = phase_ispec_inner_acoustic[q + (int64_t)np * ph];
        if (e > 0) lst.push_back(e - 1);
      }
    m->nspec_acoustic = (int)lst.size();
    m->acoustic_elems = dupload(lst.data(), (int64_t)lst.size());
    // range of the global points of acoustic elements: outside it the potentials stay zero, so the
    // Newmark updates skip it (a fluid layer on top of a solid is a small slice of NGLOB)
    int gmin = m->nglob, gmax = -1;
    for (int e : lst)
      for (int p = 0; p < 125; p++) {
        const int g = H->ibool[(int64_t)e * 125 + p];
        gmin = std::min(gmin, g);
        gmax = std::max(gmax, g);
      }
    m->ac_g0 = gmax < 0 ? 0 : (gmin / 4) * 4;
    m->ac_n = gmax < 0 ? 0 : (int64_t)gmax + 1 - m->ac_g0;
    if (m->assembly == 1) {
      build_csr(m, lst, m->asm_ac_ptr, m->asm_ac_idx);
      m->asm_ac_buf = dmalloc<realw>(nloc);
    }
    if (m->assembly == 2) build_colors(m, lst, m->ac_color_elems, m->ac_color_off, m->ac_ncolor);
  }
  m->ac[0].pot = dmalloc<realw>(n);
  m->ac[0].potd = dmalloc<realw>(n);
  m->ac[0].potdd = dmalloc<realw>(n);
  SF_CHECK(hipMemset(m->ac[0].pot, 0, sizeof(realw) * n));
  SF_CHECK(hipMemset(m->ac[0].potd, 0, sizeof(realw) * n));
  SF_CHECK(hipMemset(m->ac[0].potdd, 0, sizeof(realw) * n));
  // free surface points of acoustic elements
  {
    RawFaces rf;
    const int nf = *num_free_surface_faces;
    for (int f = 0; f < nf; f++) {
      rf.ispec.push_back(free_surface_ispec[f] - 1);
      for (int q = 0; q < 75; q++) rf.ijk.push_back(free_surface_ijk[(size_t)f * 75 + q] - 1);
    }
    build_face_points(m, rf, H->is_ac.data(), H->ibool.data(), m->free_ac, nullptr, nullptr);
  }
  // coupling faces (stored on the acoustic element)
  {
    RawFaces rf;
    const int nf = *num_coupling_ac_el_faces;
    for (int f = 0; f < nf; f++) {
      rf.ispec.push_back(coupling_ac_el_ispec[f] - 1);
      for (int q = 0; q < 75; q++) rf.ijk.push_back(coupling_ac_el_ijk[(size_t)f * 75 + q] - 1);
    }
    rf.normal.assign(coupling_ac_el_normal, coupling_ac_el_normal + (size_t)nf * 75);
    rf.jacw.assign(coupling_ac_el_jacobian2Dw, coupling_ac_el_jacobian2Dw + (size_t)nf * 25);
    build_face_points(m, rf, nullptr, H->ibool.data(), m->coupling, nullptr, nullptr);
  }
  // Stacey face points in acoustic elements: rho and c = sqrt(kappa/rho) per point
  if (m->stacey && m->num_abs_boundary_faces > 0) {
    std::vector<int> es, pl;
    build_face_points(m, H->abs, H->is_ac.data(), H->ibool.data(), m->abs_ac, &es, &pl);
    prep_log("acoustic Stacey points built");
    std::vector<realw> r(m->abs_ac.n), c(m->abs_ac.n);
    for (int i = 0; i < m->abs_ac.n; i++) {
      const int64_t q = (int64_t)es[i] * 125 + pl[i];
      r[i] = rhostore[q];
      c[i] = std::sqrt(kappastore[q] / rhostore[q]);
    }
    m->abs_ac.imp_p = dupload(r.data(), m->abs_ac.n);
    m->abs_ac.imp_s = dupload(c.data(), m->abs_ac.n);
    if (m->save_forward || m->simulation_type == 3)
      m->b_absorb_potential = dmalloc<realw>(25 * (int64_t)std::max(1, m->num_abs_boundary_faces));
  }
}

// prepare_mesh_constants_cuda.cu:1176 -- C-PML elements of the elastic domain. The convolution
// coefficients arrive in the Fortran layouts pml_convolution_coef_alpha/beta(9,NGLL3,NSPEC_CPML),
// _abar(5,..), _strain(18,..); they are regrouped into per-element blocks of 125-point planes with
// wgll_cube * rhostore appended. The stiffness launches drop the C-PML elements, which a second
// kernel computes (as compute_forces_viscoelastic_cuda does through pml_kernel_2_impl).
extern "C" void FC_FUNC_(prepare_fields_elastic_pml, PREPARE_FIELDS_ELASTIC_PML)(
    long* Mesh_pointer, int* NSPEC_CPML, int* is_CPML, int* CPML_to_spec, int* spec_to_CPML,
    realw* pml_convolution_coef_alpha, realw* pml_convolution_coef_beta, realw* pml_convolution_coef_abar,
    realw* pml_convolution_coef_strain, realw* h_wgll_cube, realw* rhostore) {
  Mesh* m = from_handle(Mesh_pointer);
  if (!m->pml_conditions) return;
  HostSide* H = host_of(m);
  const char* where = "prepare_fields_elastic_pml";
  if (!m->kappamu || !m->elastic_elems) fatal(where, "prepare_fields_elastic_device must come first");
  if (m->pml_state || m->pml_n) fatal(where, "C-PML arrays are already prepared");
  if (m->simulation_type != 1)
    fatal(where, "C-PML runs are forward only (SIMULATION_TYPE = 1), as the reference's GPU C-PML kernel");
  if (m->nspec_acoustic > 0)  // compute_forces_acoustic_calling_routine.f90:97
    fatal(where, "C-PML conditions with fluid elements are not supported on the GPU, as in the reference");
  const int ncp = *NSPEC_CPML;
  m->nspec_cpml = ncp;
  H->is_cpml.assign(m->nspec, 0);
  for (int e = 0; e < m->nspec; e++) {
    const bool a = is_CPML[e] != 0, b = spec_to_CPML[e] > 0;
    if (a != b || spec_to_CPML[e] > ncp || (b && CPML_to_spec[spec_to_CPML[e] - 1] != e + 1))
      fatal(where, "is_CPML, spec_to_CPML and CPML_to_spec disagree");
    H->is_cpml[e] = a;
  }
  for (int ic = 0; ic < ncp; ic++) {
    const int e = CPML_to_spec[ic] - 1;
    if (e < 0 || e >= m->nspec) fatal(where, "CPML_to_spec out of range");
    if (!H->is_el[e])
      fatal(where, "C-PML elements in acoustic domains are not supported (elastic C-PML only, as the reference's GPU path)");
  }
  if (ncp == 0) return;
  SF_CHECK(hipStreamSynchronize(m->stream));

  // stiffness list without the C-PML elements; C-PML list in the same processing order
  std::vector<int> lst(m->nspec_elastic), keep, pe, pc;
  SF_CHECK(hipMemcpy(lst.data(), m->elastic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
  for (int e : lst) {
    if (H->is_cpml[e]) {
      pe.push_back(e);
      pc.push_back(spec_to_CPML[e] - 1);
    } else {
      keep.push_back(e);
    }
  }
  dfree(m->elastic_elems);
  m->nspec_elastic = (int)keep.size();
  m->elastic_elems = keep.empty() ? dmalloc<int>(1) : dupload(keep.data(), (int64_t)keep.size());
  if (m->elastic_launch) {
    std::vector<int> lo(lst.size()), lk;
    SF_CHECK(hipMemcpy(lo.data(), m->elastic_launch, sizeof(int) * lo.size(), hipMemcpyDeviceToHost));
    for (int e : lo)
      if (!H->is_cpml[e]) lk.push_back(e);
    dfree(m->elastic_launch);
    m->elastic_launch = lk.empty() ? nullptr : dupload(lk.data(), (int64_t)lk.size());
  }
  m->pml_n = (int)pe.size();
  m->pml_elems = dupload(pe.data(), (int64_t)pe.size());
  m->pml_cpml = dupload(pc.data(), (int64_t)pc.size());

  // coefficient planes
  {
    const size_t blk = (size_t)SF_PML_NCOEF * 125;
    std::vector<realw> cf((size_t)ncp * blk);
    for (int ic = 0; ic < ncp; ic++) {
      const int64_t e = CPML_to_spec[ic] - 1;
      for (int p = 0; p < 125; p++) {
        realw* b = &cf[(size_t)ic * blk + p];
        const int64_t q = p + 125 * (int64_t)ic;
        for (int t = 0; t < 9; t++) {
          b[t * 125] = pml_convolution_coef_alpha[t + 9 * q];
          b[(9 + t) * 125] = pml_convolution_coef_beta[t + 9 * q];
        }
        for (int t = 0; t < 5; t++) b[(18 + t) * 125] = pml_convolution_coef_abar[t + 5 * q];
        for (int t = 0; t < 18; t++) b[(23 + t) * 125] = pml_convolution_coef_strain[t + 18 * q];
        b[41 * 125] = h_wgll_cube[p] * rhostore[e * 125 + p];
      }
    }
    // Per element: coefficient planes that hold one value (kept as that value), and state planes
    // whose multiplier plane is zero over the element (their memory variables never reach the
    // stress or the acceleration, so they are not kept). Face regions damp one direction: about a
    // third of the state and of the coefficients are live there. The rest is packed: the
    // non-uniform coefficient planes and the live state planes of an element are contiguous, in
    // plane order, at per-element offsets.
    std::vector<uint64_t> mk(2 * (size_t)ncp);
    for (int ic = 0; ic < ncp; ic++) {
      const realw* b = &cf[(size_t)ic * blk];
      uint64_t um = 0, nz = 0;
      for (int m = 0; m < SF_PML_NCOEF; m++) {
        bool uni = true, any = false;
        for (int p = 0; p < 125; p++) {
          uni = uni && b[m * 125 + p] == b[m * 125];
          any = any || b[m * 125 + p] != 0.f;
        }
        if (uni) um |= 1ull << m;
        if (any) nz |= 1ull << m;
      }
      uint64_t live = 0;
      for (int a = 0; a < 9; a++)
        for (int s2 = 0; s2 < 3; s2++)
          if (nz >> (23 + 4 * (a / 3) + 1 + s2) & 1) live |= 1ull << (a * 3 + s2);
      for (int bb = 0; bb < 12; bb++)
        if (nz >> (23 + 13 + 2 * (bb / 4)) & 1) live |= 1ull << (27 + bb);
      for (int c = 0; c < 3; c++)
        for (int s2 = 0; s2 < 3; s2++)
          if (nz >> (20 + s2) & 1) live |= 1ull << (39 + 3 * c + s2);
      mk[2 * (size_t)ic] = um;
      mk[2 * (size_t)ic + 1] = live;
    }
    std::vector<int64_t> off(2 * (size_t)ncp);
    std::vector<realw> packed;
    int64_t nst = 0;
    for (int ic = 0; ic < ncp; ic++) {
      const realw* b = &cf[(size_t)ic * blk];
      off[2 * (size_t)ic] = (int64_t)packed.size();
      off[2 * (size_t)ic + 1] = nst;
      for (int m2 = 0; m2 < SF_PML_NCOEF; m2++) {
        const int len = (mk[2 * (size_t)ic] >> m2 & 1) ? 1 : 125;
        packed.insert(packed.end(), b + m2 * 125, b + m2 * 125 + len);
      }
      packed.resize((packed.size() + 31) / 32 * 32, (realw)0);  // elements start on 128-byte lines
      nst += (125 * (int64_t)__builtin_popcountll(mk[2 * (size_t)ic + 1]) + 31) / 32 * 32;
    }
    m->pml_coef = dupload(packed.data(), (int64_t)packed.size());
    m->pml_masks = dupload(mk.data(), (int64_t)mk.size());
    m->pml_off = dupload(off.data(), (int64_t)off.size());
    m->pml_state = dmalloc<realw>(std::max<int64_t>(1, nst));
    SF_CHECK(hipMemset(m->pml_state, 0, sizeof(realw) * std::max<int64_t>(1, nst)));
  }
  m->pml_old = dmalloc<realw>(3 * (int64_t)m->nglob);
  SF_CHECK(hipMemset(m->pml_old, 0, sizeof(realw) * 3 * (size_t)m->nglob));
  // global points of the C-PML elements, and those on the outer faces of the layers
  {
    std::vector<unsigned char> mark((size_t)m->nglob, 0);
    for (int e : pe)
      for (int p = 0; p < 125; p++) mark[H->ibool[(int64_t)e * 125 + p]] = 1;
    std::vector<int> pts;
    for (int g = 0; g < m->nglob; g++)
      if (mark[g]) pts.push_back(g);
    m->pml_npts = (int)pts.size();
    m->pml_pts = dupload(pts.data(), (int64_t)pts.size());
    std::fill(mark.begin(), mark.end(), 0);
    const RawFaces& rf = H->abs;
    for (size_t f = 0; f < rf.ispec.size(); f++) {
      const int e = rf.ispec[f];
      if (!H->is_el[e] || !H->is_cpml[e]) continue;
      for (int q = 0; q < 25; q++) {
        const int* ijk = &rf.ijk[(f * 25 + q) * 3];
        mark[H->ibool[(int64_t)e * 125 + ijk[0] + 5 * ijk[1] + 25 * ijk[2]]] = 1;
      }
    }
    std::vector<int> bc;
    for (int g = 0; g < m->nglob; g++)
      if (mark[g]) bc.push_back(g);
    m->pml_nbc = (int)bc.size();
    m->pml_bc_pts = bc.empty() ? nullptr : dupload(bc.data(), (int64_t)bc.size());
  }
  // the assembly structures follow the new lists
  if (m->assembly == 1) build_elastic_assembly(m);
  if (m->assembly == 2) {
    std::vector<int> order = keep;
    if (m->elastic_launch) {
      order.resize(m->nspec_elastic);
      SF_CHECK(hipMemcpy(order.data(), m->elastic_launch, sizeof(int) * order.size(), hipMemcpyDeviceToHost));
    }
    build_colors(m, order, m->el_color_elems, m->el_color_off, m->el_ncolor);
  }
  prep_log("C-PML arrays prepared");
}

// prepare_mesh_constants_cuda.cu:999
extern "C" void FC_FUNC_(prepare_fields_elastic_adj_dev, PREPARE_FIELDS_ELASTIC_ADJ_DEV)(
    long* Mesh_pointer, int* size_f, realw* epsilon_trace_over_3, realw* b_epsilondev_xx, realw* b_epsilondev_yy,
    realw* b_epsilondev_xy, realw* b_epsilondev_xz, realw* b_epsilondev_yz, realw* b_epsilon_trace_over_3,
    int* R_size, realw* b_R_xx, realw* b_R_yy, realw* b_R_xy, realw* b_R_xz, realw* b_R_yz, realw* b_R_trace,
    realw* b_epsilondev_trace, realw* b_alphaval, realw* b_betaval, realw* b_gammaval, int* ANISOTROPIC_KL,
    int* APPROXIMATE_HESS_KL) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*ANISOTROPIC_KL) fatal("prepare_fields_elastic_adj_dev", "anisotropic kernels are not supported");
  if (*APPROXIMATE_HESS_KL) fatal("prepare_fields_elastic_adj_dev", "Hessian kernels are not supported");
  const int64_t n3 = *size_f, nloc = (int64_t)m->nspec * 125;
  m->el[1].displ = dmalloc<realw>(n3);
  m->el[1].veloc = dmalloc<realw>(n3);
  m->el[1].accel = dmalloc<realw>(n3);
  SF_CHECK(hipMemset(m->el[1].displ, 0, sizeof(realw) * n3));
  SF_CHECK(hipMemset(m->el[1].veloc, 0, sizeof(realw) * n3));
  SF_CHECK(hipMemset(m->el[1].accel, 0, sizeof(realw) * n3));
  m->eps_trace_over_3[0] = dupload(epsilon_trace_over_3, nloc);
  m->eps_trace_over_3[1] = dupload(b_epsilon_trace_over_3, nloc);
  {
    const realw* src[6] = {b_epsilondev_xx, b_epsilondev_yy, b_epsilondev_xy, b_epsilondev_xz, b_epsilondev_yz,
                           m->attenuation ? b_epsilondev_trace : nullptr};
    m->att[1].eps = upload_strain_blocks(src, nloc);
  }
  if (!m->att[0].eps) {
    m->att[0].eps = dmalloc<realw>(6 * nloc);
    SF_CHECK(hipMemset(m->att[0].eps, 0, sizeof(realw) * 6 * nloc));
  }
  if (m->attenuation) {
    if (*R_size != 3 * nloc) fatal("prepare_fields_elastic_adj_dev", "R_size must be N_SLS*125*NSPEC");
    const realw* rs[6] = {b_R_xx, b_R_yy, b_R_xy, b_R_xz, b_R_yz, b_R_trace};
    m->att[1].R = upload_memory_blocks(rs, nloc);
    for (int s = 0; s < 3; s++) {
      m->b_alphaval[s] = b_alphaval[s];
      m->b_betaval[s] = b_betaval[s];
      m->b_gammaval[s] = b_gammaval[s];
    }
  }
  m->rho_kl = dmalloc<realw>(nloc);
  m->mu_kl = dmalloc<realw>(nloc);
  m->kappa_kl = dmalloc<realw>(nloc);
  SF_CHECK(hipMemset(m->rho_kl, 0, sizeof(realw) * nloc));
  SF_CHECK(hipMemset(m->mu_kl, 0, sizeof(realw) * nloc));
  SF_CHECK(hipMemset(m->kappa_kl, 0, sizeof(realw) * nloc));
}

// prepare_mesh_constants_cuda.cu:573
extern "C" void FC_FUNC_(prepare_fields_acoustic_adj_dev, PREPARE_FIELDS_ACOUSTIC_ADJ_DEV)(long* Mesh_pointer,
                                                                                           int* APPROXIMATE_HESS_KL) {
  Mesh* m = from_handle(Mesh_pointer);
  if (*APPROXIMATE_HESS_KL) fatal("prepare_fields_acoustic_adj_dev", "Hessian kernels are not supported");
  const int64_t n = m->nglob, nloc = (int64_t)m->nspec * 125;
  m->ac[1].pot = dmalloc<realw>(n);
  m->ac[1].potd = dmalloc<realw>(n);
  m->ac[1].potdd = dmalloc<realw>(n);
  SF_CHECK(hipMemset(m->ac[1].pot, 0, sizeof(realw) * n));
  SF_CHECK(hipMemset(m->ac[1].potd, 0, sizeof(realw) * n));
  SF_CHECK(hipMemset(m->ac[1].potdd, 0, sizeof(realw) * n));
  m->rho_ac_kl = dmalloc<realw>(nloc);
  m->kappa_ac_kl = dmalloc<realw>(nloc);
  SF_CHECK(hipMemset(m->rho_ac_kl, 0, sizeof(realw) * nloc));
  SF_CHECK(hipMemset(m->kappa_ac_kl, 0, sizeof(realw) * nloc));
}

// prepare_mesh_constants_cuda.cu:1306 -- adjoint sources at receivers (SIMULATION_TYPE 2/3)
extern "C" void FC_FUNC_(prepare_sim2_or_3_const_device, PREPARE_SIM2_OR_3_CONST_DEVICE)(
    long* Mesh_pointer, int* nadj_rec_local, int* NTSTEP_BETWEEN_READ_ADJSRC, realw* hxir_adjstore,
    realw* hetar_adjstore, realw* hgammar_adjstore, int* nrec, int* h_islice_selected_rec,
    int* h_ispec_selected_rec) {
  Mesh* m = from_handle(Mesh_pointer);
  HostSide* H = host_of(m);
  m->nadj_rec_local = *nadj_rec_local;
  m->ntstep_between_read_adjsrc = *NTSTEP_BETWEEN_READ_ADJSRC;
  if (m->nadj_rec_local <= 0) return;
  // adjoint sources act in the domain of the receiver element: elastic ones on accel
  // (compute_add_sources_viscoelastic.F90:226), acoustic ones on potential_dot_dot_acoustic
  // (compute_add_sources_acoustic.f90:205)
  std::vector<int> own_el, idx_el, own_ac, idx_ac;
  std::vector<realw> h((size_t)m->nadj_rec_local * 125);
  int ir = 0;
  m->h_adj_elastic = new unsigned char[m->nadj_rec_local];

  for (int r = 0; r < *nrec && ir < m->nadj_rec_local; r++) {
    if (h_islice_selected_rec[r] != m->myrank) continue;
    const int e = h_ispec_selected_rec[r] - 1;
    if (H->is_el[e]) {
      own_el.push_back(e);
      idx_el.push_back(ir);
    } else if (H->is_ac[e]) {
      own_ac.push_back(e);
      idx_ac.push_back(ir);
    }
    m->h_adj_elastic[ir] = H->is_el[e];
    for (int p = 0; p < 125; p++) {
      const int i = p % 5, j = (p / 5) % 5, k = p / 25;
      h[(size_t)ir * 125 + p] = hxir_adjstore[i + 5 * ir] * hetar_adjstore[j + 5 * ir] * hgammar_adjstore[k + 5 * ir];
    }
    ir++;
  }
  auto remap_owner = [](FacePoints& fp, const std::vector<int>& idx) {
    std::vector<int> o(fp.n);
    SF_CHECK(hipMemcpy(o.data(), fp.ispec, sizeof(int) * o.size(), hipMemcpyDeviceToHost));
    for (auto& v : o) v = idx[v];
    SF_CHECK(hipMemcpy(fp.ispec, o.data(), sizeof(int) * o.size(), hipMemcpyHostToDevice));
  };
  if (!own_el.empty()) {
    build_point_sets(own_el, H->ibool, m->adj_pts);
    remap_owner(m->adj_pts, idx_el);
  }
  if (!own_ac.empty()) {
    build_point_sets(own_ac, H->ibool, m->adj_ac_pts);
    remap_owner(m->adj_ac_pts, idx_ac);
  }
  m->adj_h = dupload(h.data(), (int64_t)h.size());
  m->adj_src = dmalloc<realw>(3 * (int64_t)m->nadj_rec_local);
}

// prepare_mesh_constants_cuda.cu:1552
extern "C" void FC_FUNC_(prepare_cleanup_device, PREPARE_CLEANUP_DEVICE)(long* Mesh_pointer, int* ACOUSTIC_SIMULATION,
                                                                         int* ELASTIC_SIMULATION,
                                                                         int* NOISE_TOMOGRAPHY) {
  (void)ACOUSTIC_SIMULATION; (void)ELASTIC_SIMULATION; (void)NOISE_TOMOGRAPHY;
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  free_checkpoints(m);
  dfree(m->ibool); dfree(m->irregular); dfree(m->jac); dfree(m->elastic_elems); dfree(m->acoustic_elems);
  dfree(m->el_color_elems); dfree(m->ac_color_elems); dfree(m->elastic_launch);
  dfree(m->kappamu); dfree(m->rho); dfree(m->kappa_ac); dfree(m->rmass); dfree(m->rmass_ac);
  for (int f = 0; f < 2; f++) {
    dfree(m->el[f].displ); dfree(m->el[f].veloc); dfree(m->el[f].accel);
    dfree(m->ac[f].pot); dfree(m->ac[f].potd); dfree(m->ac[f].potdd);
    dfree(m->att[f].R); dfree(m->att[f].eps); dfree(m->eps_trace_over_3[f]);
  }
  dfree(m->factor_common); dfree(m->factor_common_kappa); dfree(m->fc_table); dfree(m->fc_index);
  free_face_points(m->abs_el); free_face_points(m->abs_ac); free_face_points(m->free_ac);
  free_face_points(m->coupling);
  dfree(m->b_absorb_field); dfree(m->b_absorb_potential);
  free_face_points(m->src_pts); free_face_points(m->src_ac_pts);
  dfree(m->src_array); dfree(m->src_kappa);
  free_face_points(m->adj_pts); free_face_points(m->adj_ac_pts);
  dfree(m->adj_h); dfree(m->adj_src);
  dfree(m->rec_iglob); dfree(m->rec_h); dfree(m->rec_nu);
  dfree(m->rec_ispec); dfree(m->rec_el_list); dfree(m->rec_ac_list);
  dfree(m->seismo_d); dfree(m->seismo_v); dfree(m->seismo_a); dfree(m->seismo_p);
  dfree(m->rho_kl); dfree(m->mu_kl); dfree(m->kappa_kl); dfree(m->rho_ac_kl); dfree(m->kappa_ac_kl);
  dfree(m->asm_ptr); dfree(m->asm_idx); dfree(m->asm_buf);
  dfree(m->pml_elems); dfree(m->pml_cpml);
  dfree(m->pml_state); dfree(m->pml_coef); dfree(m->pml_masks); dfree(m->pml_off);
  dfree(m->pml_old); dfree(m->pml_pts); dfree(m->pml_bc_pts);
  dfree(m->asm_ac_ptr); dfree(m->asm_ac_idx); dfree(m->asm_ac_buf);
  dfree(m->d_norm);
  if (m->h_norm) (void)hipHostFree(m->h_norm);
  if (m->timing_ev) {
    for (int i = 0; i < 2 * m->timing_max; i++) (void)hipEventDestroy(m->timing_ev[i]);
    delete[] m->timing_ev;
  }
  delete[] m->h_src_elastic; delete[] m->h_src_acoustic; delete[] m->h_rec_elastic; delete[] m->h_rec_ispec; delete[] m->h_adj_elastic;
  drop_host(m);
  SF_CHECK(hipStreamDestroy(m->stream));
  delete m;
  *Mesh_pointer = 0;
}

// ------------------------------------------------------------------------------------------------
// Library extensions (no counterpart in the reference interface).

// exact = 1: element kernels without fused multiply-add (the CPU reference's rounding);
// assembly = 1: element contributions summed per global point in element order (deterministic,
// the reference's serial order) instead of float atomics; assembly = 2: one launch per colour of
// a greedy element colouring, plain read-modify-write scatter (deterministic, no atomics).
extern "C" void sf_set_options(long* Mesh_pointer, int exact, int assembly) {
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  m->exact = exact ? 1 : 0;
  m->assembly = assembly;
  if (assembly == 1) {
    if (!m->asm_ready && m->nspec_elastic > 0 && m->kappamu) build_elastic_assembly(m);
    if (!m->asm_ac_buf && m->nspec_acoustic > 0 && m->acoustic_elems) {
      std::vector<int> lst(m->nspec_acoustic);
      SF_CHECK(hipMemcpy(lst.data(), m->acoustic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
      build_csr(m, lst, m->asm_ac_ptr, m->asm_ac_idx);
      m->asm_ac_buf = dmalloc<realw>((int64_t)m->nspec * 125);
    }
  }
  if (assembly == 2) {
    if (!m->el_color_elems) ensure_element_colors(m);
    if (!m->ac_color_elems && m->nspec_acoustic > 0 && m->acoustic_elems) {
      std::vector<int> lst(m->nspec_acoustic);
      SF_CHECK(hipMemcpy(lst.data(), m->acoustic_elems, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
      build_colors(m, lst, m->ac_color_elems, m->ac_color_off, m->ac_ncolor);
    }
  }
}

// Tuning switches read by the step calls: "att_variant" (0..4: attenuation-state scheduling of the
// point kernel, same results; 4 is the default), "acoustic_kernel_field" (below). Unknown names
// are a fatal error.
extern "C" void sf_set_tuning(long* Mesh_pointer, const char* name, int value) {
  Mesh* m = from_handle(Mesh_pointer);
  SF_CHECK(hipStreamSynchronize(m->stream));
  if (std::strcmp(name, "att_variant") == 0) {
    if (value < 0 || value > 4) fatal("sf_set_tuning", "att_variant must be 0..4");
    m->att_variant = value;
  } else if (std::strcmp(name, "acoustic_kernel_field") == 0) {
    // which adjoint array compute_kernels_acoustic_cuda pairs with the reconstructed field:
    // 0 = potential_acoustic (compute_kernels.f90:269 and :325), 1 = potential_dot_dot_acoustic
    if (value < 0 || value > 1) fatal("sf_set_tuning", "acoustic_kernel_field must be 0 or 1");
    m->ac_kernel_dd = value;
  } else if (std::strcmp(name, "timing_scope") == 0) {
    // what sf_timing_start/sf_timing_read time in each elastic stiffness call: 0 = the whole call,
    // 1 = its C-PML launch alone (calls without one are not counted)
    if (value < 0 || value > 1) fatal("sf_set_tuning", "timing_scope must be 0 or 1");
    m->timing_scope = value;
  } else {
    fatal("sf_set_tuning", "unknown tuning switch");
  }
}

// Launch order of the elastic elements for the atomic and coloured stiffness launches (results do
// not depend on it except for float-atomic rounding): `order` lists the 0-based elastic elements,
// each once. Hosts pass a locality order (e.g. Morton order of element centres) so that elements
// sharing points run close together and their shared rows meet in L2. Re-colours when coloured.
extern "C" void sf_set_element_order(long* Mesh_pointer, int n, const int* order) {
  Mesh* m = from_handle(Mesh_pointer);
  HostSide* H = host_of(m);
  // with C-PML the order may list the C-PML elements too (they are launched separately)
  if (n != m->nspec_elastic && n != m->nspec_elastic + m->pml_n)
    fatal("sf_set_element_order", "the order must list every elastic element once");
  std::vector<unsigned char> seen((size_t)m->nspec, 0);
  std::vector<int> lst;
  lst.reserve(n);
  for (int q = 0; q < n; q++) {
    const int e = order[q];
    if (e < 0 || e >= m->nspec || !H->is_el[e] || seen[e]) fatal("sf_set_element_order", "not a permutation of the elastic elements");
    seen[e] = 1;
    if (H->is_cpml.empty() || !H->is_cpml[e]) lst.push_back(e);
  }
  if ((int)lst.size() != m->nspec_elastic) fatal("sf_set_element_order", "not a permutation of the elastic elements");
  SF_CHECK(hipStreamSynchronize(m->stream));
  dfree(m->elastic_launch);
  m->elastic_launch = dupload(lst.data(), (int64_t)lst.size());
  if (m->el_color_elems) build_colors(m, lst, m->el_color_elems, m->el_color_off, m->el_ncolor);
}

// Greedy element colouring used by assembly = 2 (host only, no device needed): colour[q] of
// elems[q] (0-based element ids into ibool0, 0-based, 125 per element); returns the number of
// colours, or -1 if more than 32 would be needed.
extern "C" int sf_color_elements(int nglob, int nelem, const int* ibool0, const int* elems, int* color) {
  return color_elements(nglob, nelem, ibool0, elems, color);
}

// Replace the recomputed inverse determinant of irregular elements by the host's jacobianstore
// (5,5,5,NSPEC_IRREGULAR), so the irregular path uses the CPU solver's exact values.
extern "C" void sf_set_jacobian(long* Mesh_pointer, const realw* jacobianstore) {
  Mesh* m = from_handle(Mesh_pointer);
  const int64_t P = (int64_t)m->nspec_irregular * 125;
  if (P <= 0) return;
  SF_CHECK(hipMemcpy(m->jac + 9 * P, jacobianstore, sizeof(realw) * (size_t)P, hipMemcpyHostToDevice));
  m->jac_from_host = 1;
}
