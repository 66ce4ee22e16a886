/*
 * specfem3d_amd.h -- C-ABI of libspecfem3d_amd.so, the MI355X (gfx950) replacement of
 * SPECFEM3D Cartesian's GPU interface for the single-partition (NPROC=1) solver and adjoint loops.
 *
 * The symbols are the Fortran-mangled names (lower case + trailing underscore) the reference's
 * Fortran host already calls through implicit interfaces; every argument is passed by reference,
 * Fortran logicals arrive as int, arrays are column-major with 1-based index values, `realw` is
 * float. Each declaration names the reference interface it replaces (the authoritative signature
 * list is src/gpu/specfem3D_gpu_cuda_method_stubs.c). The sf_* functions at the end have no
 * counterpart in the reference.
 */
#ifndef SPECFEM3D_AMD_H
#define SPECFEM3D_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef float realw;

/* ---- device and set-up ------------------------------------------------------------------- */
/* src/gpu/initialize_gpu.cu:535 */
void initialize_gpu_device_(int* myrank_f, int* ncuda_devices);
/* src/gpu/prepare_mesh_constants_cuda.cu:70 (stubs.c:481) */
void prepare_constants_device_(long* Mesh_pointer, int* h_NGLLX, int* NSPEC_AB, int* NGLOB_AB, int* NSPEC_IRREGULAR,
                               int* h_irregular_element_number, realw* h_xix, realw* h_xiy, realw* h_xiz,
                               realw* h_etax, realw* h_etay, realw* h_etaz, realw* h_gammax, realw* h_gammay,
                               realw* h_gammaz, realw* xix_regular, realw* jacobian_regular, int* h_ibool,
                               int* num_interfaces_ext_mesh, int* max_nibool_interfaces_ext_mesh,
                               int* h_nibool_interfaces_ext_mesh, int* h_ibool_interfaces_ext_mesh,
                               realw* h_hprime_xx, realw* h_hprimewgll_xx, realw* h_wgllwgll_xy,
                               realw* h_wgllwgll_xz, realw* h_wgllwgll_yz, int* STACEY_ABSORBING_CONDITIONS,
                               int* h_abs_boundary_ispec, int* h_abs_boundary_ijk, realw* h_abs_boundary_normal,
                               realw* h_abs_boundary_jacobian2Dw, int* h_num_abs_boundary_faces,
                               int* h_ispec_is_inner, int* NSOURCES, int* nsources_local_f, realw* h_sourcearrays,
                               int* h_islice_selected_source, int* h_ispec_selected_source,
                               int* h_ispec_selected_rec, int* nrec, int* nrec_local, int* SIMULATION_TYPE,
                               int* USE_MESH_COLORING_GPU_f, int* nspec_acoustic, int* nspec_elastic,
                               int* ispec_is_acoustic, int* ispec_is_elastic, int* h_myrank, int* SAVE_FORWARD,
                               realw* h_xir, realw* h_etar, realw* h_gammar, double* nu_rec, double* nu_source,
                               int* h_islice_selected_rec, int* nlength_seismogram,
                               int* SAVE_SEISMOGRAMS_DISPLACEMENT, int* SAVE_SEISMOGRAMS_VELOCITY,
                               int* SAVE_SEISMOGRAMS_ACCELERATION, int* SAVE_SEISMOGRAMS_PRESSURE,
                               int* h_NB_RUNS_ACOUSTIC_GPU, int* FAULT_SIMULATION,
                               int* UNDO_ATTENUATION_AND_OR_PML, int* PML_CONDITIONS);
/* src/gpu/prepare_mesh_constants_cuda.cu:657 (stubs.c:546) */
void prepare_fields_elastic_device_(long* Mesh_pointer, realw* rmassx, realw* rmassy, realw* rmassz, realw* rho_vp,
                                    realw* rho_vs, realw* h_kappav, realw* h_muv, int* num_phase_ispec_elastic,
                                    int* phase_ispec_inner_elastic, realw* b_absorb_field, int* b_reclen_field,
                                    int* COMPUTE_AND_STORE_STRAIN, realw* epsilondev_xx, realw* epsilondev_yy,
                                    realw* epsilondev_xy, realw* epsilondev_xz, realw* epsilondev_yz,
                                    int* ATTENUATION, int* R_size, realw* R_xx, realw* R_yy, realw* R_xy, realw* R_xz,
                                    realw* R_yz, realw* factor_common, realw* R_trace, realw* epsilondev_trace,
                                    realw* factor_common_kappa, realw* alphaval, realw* betaval, realw* gammaval,
                                    int* APPROXIMATE_OCEAN_LOAD, realw* rmass_ocean_load, int* NOISE_TOMOGRAPHY,
                                    realw* free_surface_normal, int* free_surface_ispec, int* free_surface_ijk,
                                    int* num_free_surface_faces, int* ACOUSTIC_SIMULATION,
                                    int* num_colors_outer_elastic, int* num_colors_inner_elastic,
                                    int* num_elem_colors_elastic, int* ANISOTROPY, realw* c11store, realw* c12store,
                                    realw* c13store, realw* c14store, realw* c15store, realw* c16store,
                                    realw* c22store, realw* c23store, realw* c24store, realw* c25store,
                                    realw* c26store, realw* c33store, realw* c34store, realw* c35store,
                                    realw* c36store, realw* c44store, realw* c45store, realw* c46store,
                                    realw* c55store, realw* c56store, realw* c66store);
/* src/gpu/prepare_mesh_constants_cuda.cu:434 (stubs.c:523) */
void prepare_fields_acoustic_device_(long* Mesh_pointer, realw* rmass_acoustic, realw* rhostore, realw* kappastore,
                                     int* num_phase_ispec_acoustic, int* phase_ispec_inner_acoustic,
                                     int* NOISE_TOMOGRAPHY, int* num_free_surface_faces, int* free_surface_ispec,
                                     int* free_surface_ijk, int* b_reclen_potential, realw* b_absorb_potential,
                                     int* ELASTIC_SIMULATION, int* num_coupling_ac_el_faces,
                                     int* coupling_ac_el_ispec, int* coupling_ac_el_ijk,
                                     realw* coupling_ac_el_normal, realw* coupling_ac_el_jacobian2Dw,
                                     int* num_colors_outer_acoustic, int* num_colors_inner_acoustic,
                                     int* num_elem_colors_acoustic);
/* src/gpu/prepare_mesh_constants_cuda.cu:999 (stubs.c:584) */
void prepare_fields_elastic_adj_dev_(long* Mesh_pointer, int* size_f, realw* epsilon_trace_over_3,
                                     realw* b_epsilondev_xx, realw* b_epsilondev_yy, realw* b_epsilondev_xy,
                                     realw* b_epsilondev_xz, realw* b_epsilondev_yz, realw* b_epsilon_trace_over_3,
                                     int* R_size, realw* b_R_xx, realw* b_R_yy, realw* b_R_xy, realw* b_R_xz,
                                     realw* b_R_yz, realw* b_R_trace, realw* b_epsilondev_trace, realw* b_alphaval,
                                     realw* b_betaval, realw* b_gammaval, int* ANISOTROPIC_KL,
                                     int* APPROXIMATE_HESS_KL);
/* src/gpu/prepare_mesh_constants_cuda.cu:573 (stubs.c:542) */
void prepare_fields_acoustic_adj_dev_(long* Mesh_pointer, int* APPROXIMATE_HESS_KL);
/* src/gpu/prepare_mesh_constants_cuda.cu:1306 (stubs.c:611) */
void prepare_sim2_or_3_const_device_(long* Mesh_pointer, int* nadj_rec_local, int* NTSTEP_BETWEEN_READ_ADJSRC,
                                     realw* hxir_adjstore, realw* hetar_adjstore, realw* hgammar_adjstore,
                                     int* nrec, int* h_islice_selected_rec, int* h_ispec_selected_rec);
/* src/gpu/prepare_mesh_constants_cuda.cu:1552 (stubs.c:642) */
void prepare_cleanup_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION,
                             int* NOISE_TOMOGRAPHY);

/* ---- time marching ------------------------------------------------------------------------- */
/* src/gpu/update_displacement_cuda.cu:39 (stubs.c:895) */
void update_displacement_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatover2_F, realw* deltatsqover2_F,
                               int* FORWARD_OR_ADJOINT);
/* src/gpu/update_displacement_cuda.cu:192 (stubs.c:902) */
void update_displacement_ac_cuda_(long* Mesh_pointer, realw* deltat_F, realw* deltatover2_F, realw* deltatsqover2_F,
                                  int* FORWARD_OR_ADJOINT);
/* src/gpu/update_displacement_cuda.cu:287 (stubs.c:909) */
void kernel_3_a_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F, int* APPROXIMATE_OCEAN_LOAD,
                      int* FORWARD_OR_ADJOINT);
/* src/gpu/update_displacement_cuda.cu (stubs.c:916) -- ocean load only: stops with a message */
void kernel_3_b_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F, int* FORWARD_OR_ADJOINT);
/* src/gpu/update_displacement_cuda.cu:448 (stubs.c:922) */
void kernel_3_acoustic_cuda_(long* Mesh_pointer, realw* deltatover2_F, realw* b_deltatover2_F,
                             int* FORWARD_OR_ADJOINT_f);

/* ---- forces and boundary terms ---------------------------------------------------------- */
/* src/gpu/compute_forces_viscoelastic_cuda.cu:1115 (stubs.c:243) */
void compute_forces_viscoelastic_cuda_(long* Mesh_pointer, int* iphase, realw* deltat, int* nspec_outer_elastic,
                                       int* nspec_inner_elastic, int* COMPUTE_AND_STORE_STRAIN, int* ATTENUATION,
                                       int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_forces_acoustic_cuda.cu:238 (stubs.c:226) */
void compute_forces_acoustic_cuda_(long* Mesh_pointer, int* iphase, int* nspec_outer_acoustic,
                                   int* nspec_inner_acoustic, int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_forces_acoustic_cuda.cu:334 (stubs.c:233) */
void acoustic_enforce_free_surf_cuda_(long* Mesh_pointer, int* ABSORB_INSTEAD_OF_FREE_SURFACE,
                                      int* FORWARD_OR_ADJOINT);
/* src/gpu/compute_stacey_viscoelastic_cuda.cu:39 (stubs.c:298) */
void compute_stacey_viscoelastic_cuda_(long* Mesh_pointer, int* iphasef, realw* b_absorb_field,
                                       int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_stacey_viscoelastic_cuda.cu:224 (stubs.c:304) */
void compute_stacey_viscoelastic_undoatt_cuda_(long* Mesh_pointer, int* iphasef, int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_stacey_acoustic_cuda.cu (stubs.c:282) */
void compute_stacey_acoustic_cuda_(long* Mesh_pointer, int* iphasef, realw* h_b_absorb_potential,
                                   int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_stacey_acoustic_cuda.cu (stubs.c:288) */
void compute_stacey_acoustic_undoatt_cuda_(long* Mesh_pointer, int* iphasef, int* FORWARD_OR_ADJOINT_f);
/* src/gpu/compute_coupling_cuda.cu:39 (stubs.c:205) */
void compute_coupling_ac_el_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT);
/* src/gpu/compute_coupling_cuda.cu:135 (stubs.c:211) */
void compute_coupling_el_ac_cuda_(long* Mesh_pointer, int* iphasef, int* num_coupling_ac_el_facesf,
                                  int* FORWARD_OR_ADJOINT);

/* ---- sources and receivers ----------------------------------------------------------------- */
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu:39 (stubs.c:176) */
void compute_add_sources_el_cuda_(long* Mesh_pointer, double* h_stf_pre_compute, int* h_NSOURCES);
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu (stubs.c:181) */
void compute_add_sources_el_s3_cuda_(long* Mesh_pointer, double* h_stf_pre_compute, int* h_NSOURCES);
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu:224 (stubs.c:192) */
void add_sources_el_sim_type_2_or_3_(long* Mesh_pointer, realw* h_source_adjoint, int* nrec, int* nadj_rec_local,
                                     int* NTSTEP_BETWEEN_READ_ADJSRC, int* it);
/* src/gpu/compute_add_sources_acoustic_cuda.cu (stubs.c:153-170): sources/adjoint sources inside
   fluid elements stop with a message */
void compute_add_sources_ac_cuda_(long* Mesh_pointer, int* NSOURCESf, double* h_stf_pre_compute,
                                  int* run_number_of_the_source);
void compute_add_sources_ac_s3_cuda_(long* Mesh_pointer, int* NSOURCESf, double* h_stf_pre_compute,
                                     int* run_number_of_the_source);
void add_sources_ac_sim_2_or_3_cuda_(long* Mesh_pointer, realw* h_source_adjoint, int* nrec, int* nadj_rec_local,
                                     int* NTSTEP_BETWEEN_READ_ADJSRC, int* it);
/* src/gpu/write_seismograms_cuda.cu:34 (stubs.c:933) */
void compute_seismograms_cuda_(long* Mesh_pointer_f, realw* seismograms_d, realw* seismograms_v,
                               realw* seismograms_a, realw* seismograms_p, int* seismo_currentf,
                               int* nlength_seismogramf, int* itf, int* it_endf, int* ACOUSTIC_SIMULATION,
                               int* ELASTIC_SIMULATION, int* USE_TRICK_FOR_BETTER_PRESSURE);

/* ---- sensitivity kernels, strain, stability norms ----------------------------------------- */
/* src/gpu/compute_kernels_cuda.cu:39 (stubs.c:258) */
void compute_kernels_elastic_cuda_(long* Mesh_pointer, realw* deltat_f);
/* src/gpu/compute_kernels_cuda.cu (stubs.c:267): stops with a message (fluid kernels) */
void compute_kernels_acoustic_cuda_(long* Mesh_pointer, realw* deltat_f);
/* src/gpu/compute_strain_cuda.cu (stubs.c:314) */
void compute_strain_cuda_(long* Mesh_pointer, int* FORWARD_OR_ADJOINT);
/* src/gpu/check_fields_cuda.cu:375-470 (stubs.c:110-116) */
void get_norm_elastic_from_device_(realw* norm, long* Mesh_pointer, int* type);
void get_norm_acoustic_from_device_(realw* norm, long* Mesh_pointer, int* sim_type);
/* src/gpu/check_fields_cuda.cu (stubs.c:104-108) */
void get_free_device_memory_(realw* free, realw* used, realw* total);
void output_free_device_memory_(int* myrank);

/* ---- transfers (src/gpu/transfer_fields_cuda.cu:45-863, stubs.c:683-890) --------------- */
void transfer_fields_el_to_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer);
void transfer_fields_el_from_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer);
void transfer_b_fields_to_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer);
void transfer_b_fields_from_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer);
void transfer_displ_to_device_(int* size, realw* displ, long* Mesh_pointer);
void transfer_displ_from_device_(int* size, realw* displ, long* Mesh_pointer);
void transfer_veloc_to_device_(int* size, realw* veloc, long* Mesh_pointer);
void transfer_veloc_from_device_(int* size, realw* veloc, long* Mesh_pointer);
void transfer_accel_to_device_(int* size, realw* accel, long* Mesh_pointer);
void transfer_accel_from_device_(int* size, realw* accel, long* Mesh_pointer);
void transfer_b_displ_to_device_(int* size, realw* b_displ, long* Mesh_pointer);
void transfer_b_displ_from_device_(int* size, realw* b_displ, long* Mesh_pointer);
void transfer_b_veloc_to_device_(int* size, realw* b_veloc, long* Mesh_pointer);
void transfer_b_veloc_from_device_(int* size, realw* b_veloc, long* Mesh_pointer);
void transfer_b_accel_to_device_(int* size, realw* b_accel, long* Mesh_pointer);
void transfer_b_accel_from_device_(int* size, realw* b_accel, long* Mesh_pointer);
void transfer_fields_ac_to_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                   realw* potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_fields_ac_from_device_(int* size, realw* potential_acoustic, realw* potential_dot_acoustic,
                                     realw* potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_b_fields_ac_to_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                     realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_b_fields_ac_from_device_(int* size, realw* b_potential_acoustic, realw* b_potential_dot_acoustic,
                                       realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_potential_ac_from_device_(int* size, realw* potential_acoustic, long* Mesh_pointer);
void transfer_b_potential_ac_from_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer);
void transfer_b_potential_ac_to_device_(int* size, realw* b_potential_acoustic, long* Mesh_pointer);
void transfer_b_potential_dot_dot_ac_from_device_(int* size, realw* b_potential_dot_dot_acoustic,
                                                  long* Mesh_pointer);
void transfer_b_potential_dot_dot_ac_to_device_(int* size, realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_dot_dot_to_device_(int* size, realw* potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_dot_dot_from_device_(int* size, realw* potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_b_dot_dot_from_device_(int* size, realw* b_potential_dot_dot_acoustic, long* Mesh_pointer);
void transfer_kernels_el_to_host_(long* Mesh_pointer, realw* h_rho_kl, realw* h_mu_kl, realw* h_kappa_kl,
                                  realw* h_cijkl_kl, int* NSPEC_AB);
void transfer_kernels_ac_to_host_(long* Mesh_pointer, realw* h_rho_ac_kl, realw* h_kappa_ac_kl, int* NSPEC_AB);
void transfer_strain_from_device_(long* Mesh_pointer, realw* epsilondev_xx, realw* epsilondev_yy,
                                  realw* epsilondev_xy, realw* epsilondev_xz, realw* epsilondev_yz,
                                  realw* epsilondev_trace, int* size_epsilondev);
void transfer_b_strain_to_device_(long* Mesh_pointer, realw* b_epsilondev_xx, realw* b_epsilondev_yy,
                                  realw* b_epsilondev_xy, realw* b_epsilondev_xz, realw* b_epsilondev_yz,
                                  realw* b_epsilondev_trace, int* size_epsilondev);
void transfer_rmemory_from_device_(long* Mesh_pointer, realw* R_xx, realw* R_yy, realw* R_xy, realw* R_xz,
                                   realw* R_yz, realw* R_trace, int* size_R);
void transfer_b_rmemory_to_device_(long* Mesh_pointer, realw* b_R_xx, realw* b_R_yy, realw* b_R_xy, realw* b_R_xz,
                                   realw* b_R_yz, realw* b_R_trace, int* size_R);
/* no-ops at NPROC=1 (src/gpu/assemble_MPI_*_cuda.cu, stubs.c:46-98) */
void transfer_boun_pot_from_device_(long* Mesh_pointer, realw* send_potential_dot_dot_buffer,
                                    const int* FORWARD_OR_ADJOINT);
void transfer_asmbl_pot_to_device_(long* Mesh_pointer, realw* buffer_recv_scalar_ext_mesh,
                                   const int* FORWARD_OR_ADJOINT);
void transfer_boun_accel_from_device_(long* Mesh_pointer, realw* send_accel_buffer, const int* FORWARD_OR_ADJOINT);
void transfer_boundary_from_device_a_(long* Mesh_pointer, const int* nspec_outer_elastic);
void transfer_boundary_to_device_a_(long* Mesh_pointer, realw* buffer_recv_vector_ext_mesh,
                                    const int* num_interfaces_ext_mesh, const int* max_nibool_interfaces_ext_mesh);
void transfer_asmbl_accel_to_device_(long* Mesh_pointer, realw* buffer_recv_vector_ext_mesh,
                                     const int* num_interfaces_ext_mesh, const int* max_nibool_interfaces_ext_mesh,
                                     const int* nibool_interfaces_ext_mesh, const int* ibool_interfaces_ext_mesh,
                                     const int* FORWARD_OR_ADJOINT);
void transfer_sync_accel_to_device_(long* Mesh_pointer, realw* buffer_recv_vector_ext_mesh,
                                    const int* num_interfaces_ext_mesh, const int* max_nibool_interfaces_ext_mesh,
                                    const int* nibool_interfaces_ext_mesh, const int* ibool_interfaces_ext_mesh,
                                    const int* FORWARD_OR_ADJOINT);
void sync_copy_from_device_(long* Mesh_pointer, int* iphase, realw* send_buffer);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:884-890): nothing is pinned; the library stages copies */
void register_host_array_(int* size, realw* h_array);
void unregister_host_array_(realw* h_array);

/* ---- C-PML absorbing layers (elastic elements, forward runs) ---- */
/* prepare_mesh_constants_cuda.cu:1176 -- C-PML element lists, convolution coefficients
   pml_convolution_coef_alpha/beta(9,NGLL3,NSPEC_CPML), _abar(5,..), _strain(18,..) as computed by
   prepare_timerun_pml (prepare_timerun.F90:592-905), wgll_cube(NGLL3) and rhostore(NGLL3,NSPEC_AB);
   call after prepare_fields_elastic_device_ (PML_CONDITIONS set in prepare_constants_device_) */
void prepare_fields_elastic_pml_(long* Mesh_pointer, int* NSPEC_CPML, int* is_CPML, int* CPML_to_spec,
                                 int* spec_to_CPML, realw* pml_convolution_coef_alpha,
                                 realw* pml_convolution_coef_beta, realw* pml_convolution_coef_abar,
                                 realw* pml_convolution_coef_strain, realw* h_wgll_cube, realw* rhostore);
/* pml_compute_accel_cuda.cu:245 -- Dirichlet condition on the outer faces of the layers */
void pml_impose_boundary_condition_elastic_cuda_(long* Mesh_pointer);
/* transfer_fields_cuda.cu (stubs.c:735-739) -- PML_displ_old/new(NDIM,NGLL3,NSPEC_CPML) */
void transfer_pml_displ_from_device_(int* size, realw* PML_displ_old, realw* PML_displ_new, long* Mesh_pointer);
void transfer_pml_displ_to_device_(int* size, realw* PML_displ_old, realw* PML_displ_new, long* Mesh_pointer);

/* ---- outside the single-partition hot path: these stop the run with a message (noise
   tomography, faults, gravity, ocean load, Hessian kernels, kernel smoothing, debug checks),
   as the reference's own stubs do when it is built without a GPU; pause_for_debug does nothing.
   Parameter names are placeholders. */
/* src/gpu/compute_add_sources_viscoelastic_cuda.cu (stubs.c:186) */
void add_source_main_rec_noise_cu_(long* p1, int* p2, int* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:145) */
void check_error_vectors_(int* p1, realw* p2, realw* p3);
/* src/gpu/check_fields_cuda.cu (stubs.c:142) */
void check_max_norm_b_accel_(int* p1, realw* p2, int* p3);
/* src/gpu/check_fields_cuda.cu (stubs.c:133) */
void check_max_norm_b_accel_gpu_(int* p1, realw* p2, long* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:139) */
void check_max_norm_b_displ_(int* p1, realw* p2, int* p3);
/* src/gpu/check_fields_cuda.cu (stubs.c:130) */
void check_max_norm_b_displ_gpu_(int* p1, realw* p2, long* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:136) */
void check_max_norm_b_veloc_gpu_(int* p1, realw* p2, long* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:127) */
void check_max_norm_displ_(int* p1, realw* p2, int* p3);
/* src/gpu/check_fields_cuda.cu (stubs.c:121) */
void check_max_norm_displ_gpu_(int* p1, realw* p2, long* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:124) */
void check_max_norm_vector_(int* p1, realw* p2, int* p3);
/* src/gpu/compute_coupling_cuda.cu (stubs.c:217) */
void compute_coupling_ocean_cuda_(long* p1, int* p2);
/* src/gpu/compute_kernels_cuda.cu (stubs.c:271) */
void compute_kernels_hess_cuda_(long* p1, realw* p2, int* p3, int* p4);
/* src/gpu/compute_kernels_cuda.cu (stubs.c:262) */
void compute_kernels_strgth_noise_cu_(long* p1, realw* p2, realw* p3);
/* src/gpu/smooth_cuda.cu (stubs.c:665) */
void compute_smooth_gpu_(long* p1, realw* p2, realw* p3, realw* p4, realw* p5, realw* p6, int* p7);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:427) */
void fault_solver_gpu_(long* p1, long* p2, realw* p3, int* p4);
/* src/gpu/check_fields_cuda.cu (stubs.c:118) */
void get_max_accel_(int* p1, int* p2, long* p3);
/* src/gpu/smooth_cuda.cu (stubs.c:674) */
void get_smooth_gpu_(long* p1, realw* p2);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:330) */
void initialize_fault_data_gpu_(long* p1, int* p2, int* p3, int* p4, int* p5, int* p6);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:323) */
void initialize_fault_solver_gpu_(long* p1, int* p2, realw* p3, realw* p4, int* p5);
/* src/gpu/noise_tomography_cuda.cu (stubs.c:457) */
void make_displ_rand_(long* p1, realw* p2);
/* src/gpu/noise_tomography_cuda.cu (stubs.c:463) */
void noise_read_add_surface_movie_cu_(long* p1, realw* p2, int* p3);
/* src/gpu/helper_functions.cu (stubs.c:438) */
void pause_for_debug_(void);
/* src/gpu/prepare_mesh_constants_cuda.cu (stubs.c:637) */
void prepare_fault_device_(long* p1, int* p2, realw* p3);
/* src/gpu/prepare_mesh_constants_cuda.cu (stubs.c:629) */
void prepare_fields_gravity_device_(long* p1, int* p2, realw* p3, realw* p4, realw* p5, realw* p6);
/* src/gpu/prepare_mesh_constants_cuda.cu (stubs.c:616) */
void prepare_fields_noise_device_(long* p1, int* p2, int* p3, int* p4, int* p5, int* p6, int* p7, int* p8, realw* p9, realw* p10, realw* p11, realw* p12, realw* p13, realw* p14);
/* src/gpu/smooth_cuda.cu (stubs.c:653) */
void prepare_smooth_gpu_(long* p1, realw* p2, realw* p3, realw* p4, realw* p5, realw* p6, realw* p7, realw* p8, int* p9, int* p10);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:861) */
void transfer_compute_kernel_answers_from_device_(long* p1, realw* p2, int* p3, realw* p4, int* p5, realw* p6, int* p7);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:867) */
void transfer_compute_kernel_fields_from_device_(long* p1, realw* p2, int* p3, realw* p4, int* p5, realw* p6, realw* p7, realw* p8, realw* p9, realw* p10, int* p11, realw* p12, realw* p13, realw* p14, realw* p15, realw* p16, int* p17, realw* p18, int* p19, realw* p20, int* p21, realw* p22, int* p23, realw* p24, realw* p25, int* p26);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:365) */
void transfer_datat_to_host_(long* p1, int* p2, realw* p3, int* p4);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:338) */
void transfer_fault_data_to_device_(long* p1, int* p2, int* p3, int* p4, realw* p5, realw* p6, realw* p7, realw* p8, realw* p9, realw* p10, realw* p11, realw* p12, realw* p13, int* p14, int* p15, int* p16);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:356) */
void transfer_fault_data_to_host_(long* p1, int* p2, int* p3, int* p4, realw* p5, realw* p6, realw* p7);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:854) */
void transfer_kernels_hess_ac_tohost_(long* p1, realw* p2, realw* p3, realw* p4, int* p5);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:846) */
void transfer_kernels_hess_el_tohost_(long* p1, realw* p2, realw* p3, realw* p4, realw* p5, int* p6);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:776) */
void transfer_kernels_noise_to_host_(long* p1, realw* p2, int* p3);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:371) */
void transfer_rsf_data_todevice_(long* p1, int* p2, int* p3, realw* p4, realw* p5, realw* p6, realw* p7, realw* p8, realw* p9, realw* p10, realw* p11, realw* p12, realw* p13, realw* p14, realw* p15, int* p16);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:400) */
void transfer_rsf_data_tohost_(long* p1, int* p2, int* p3, realw* p4, realw* p5, realw* p6, realw* p7, realw* p8, realw* p9, realw* p10, realw* p11, realw* p12, realw* p13, realw* p14);
/* src/gpu/transfer_fields_cuda.cu (stubs.c:720) */
void transfer_sigma_from_device_(int* p1, realw* p2, long* p3);
/* src/gpu/noise_tomography_cuda.cu (stubs.c:459) */
void transfer_surface_to_host_(long* p1, realw* p2);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:389) */
void transfer_swf_data_todevice_(long* p1, int* p2, int* p3, realw* p4, realw* p5, realw* p6, realw* p7, realw* p8, realw* p9);
/* src/gpu/fault_solver_dynamics.cu (stubs.c:416) */
void transfer_swf_data_tohost_(long* p1, int* p2, int* p3, realw* p4, realw* p5, realw* p6, realw* p7, realw* p8, realw* p9);

/* ---- library extensions (no counterpart in the reference interface) ---------------------- */
/* exact = 1: element kernels without fused multiply-add (the CPU solver's rounding);
   assembly = 0: float atomic scatter-add; 1: element contributions summed per point in element
   order (deterministic, the CPU solver's order); 2: one launch per colour of a greedy element
   colouring with a plain read-modify-write scatter (deterministic, no atomics). */
void sf_set_options(long* Mesh_pointer, int exact, int assembly);
/* tuning switch by name: "att_variant" 0..4 = scheduling of the attenuation state in the point
   stiffness kernel (same results; 4, the default, loads it straight into LDS);
   "acoustic_kernel_field" 0 = compute_kernels_acoustic_cuda pairs the adjoint potential with the
   reconstructed field as the reference does (compute_kernels.f90:269,325), 1 = it uses the adjoint
   potential_dot_dot instead (the misfit gradient when the adjoint field enters the fluid through
   the coupling); "timing_scope" 0 = sf_timing_* time each elastic stiffness call, 1 = only its
   C-PML launch; unknown names are fatal */
void sf_set_tuning(long* Mesh_pointer, const char* name, int value);
/* device-resident checkpoints for checkpointed kernel runs (the files of
   save_forward_arrays_undoatt / read_forward_arrays_undoatt, save_forward_arrays.f90:102-179 and
   read_forward_arrays.f90:133-229, and the b_displ buffer of iterate_time_undoatt.F90:45-47, kept in
   HBM): FORWARD_OR_ADJOINT 1 = forward/adjoint field, 3 = backward/reconstructed field */
void sf_snapshot_save(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int slot);
void sf_snapshot_load(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int slot);
void sf_history_reserve(long* Mesh_pointer, int nframes);
void sf_history_store(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int frame);
void sf_history_load(long* Mesh_pointer, int FORWARD_OR_ADJOINT, int frame);
void sf_checkpoints_free(long* Mesh_pointer);
/* zero the sensitivity-kernel accumulators (prepare_fields_*_adj_dev zero them once) */
void sf_zero_kernels(long* Mesh_pointer);
/* batched many-shot forward engine (no counterpart in the reference, which runs one shot per
   process): shots with their own sources, receivers and record lengths stepped together on one
   prepared mesh (isotropic elastic, optionally 3-SLS viscoelastic: each shot then carries its own
   memory variables; fluid elements coupled to the solid: each shot then carries its own
   potentials) in shot-interleaved wavefields; sources and receivers lie in elastic elements;
   element ids 0-based */
long sf_batch_create(long* Mesh_pointer, int nshots, const int* nstep, realw deltat, realw deltatover2,
                     realw deltatsqover2);
void sf_batch_add_source(long batch, int shot, int ispec, const realw* sourcearray, const realw* stf);
void sf_batch_add_receiver(long batch, int shot, int ispec, const double* hxi, const double* heta,
                           const double* hgamma, const double* nu);
void sf_batch_finalize(long batch);
/* shots per workgroup 4 or 8; assembly 0 = atomic scatter, 2 = one launch per element colour,
   3 = coloured with the Newmark update fused into the colour launches (same results as 2); only
   at step 1 (a new batch or after sf_batch_reset) */
void sf_batch_set_options(long batch, int shots_per_group, int assembly);
void sf_batch_get_options(long batch, int* shots_per_group, int* assembly);
/* kernel variants with the same results: "lds_overlay" 0/1 (displacement tile and stress products
   share LDS); unknown names are fatal */
void sf_batch_set_tuning(long batch, const char* name, int value);
int sf_batch_get_tuning(long batch, const char* name);
double sf_batch_autotune(long batch, int trial_steps);
void sf_batch_run(long batch, int nsteps, int graph_steps);
int sf_batch_current_step(long batch);
void sf_batch_seismograms(long batch, int shot, realw* out);
void sf_batch_displ(long batch, int shot, realw* out);
/* potential of one shot at every global point (zero outside the fluid elements) */
void sf_batch_potential(long batch, int shot, realw* out);
/* No counterpart in the reference (its get_norm_*_from_device allocate a partial-maximum buffer on
   every call); for callers that cannot allocate (code under HIP graph capture, serving loops): the
   stability norm of get_norm_elastic_from_device_ / get_norm_acoustic_from_device_ (largest |u|
   for ncomp = 3, |chi| for ncomp = 1, NaN counted as +inf) of the device field `field` (npts points)
   into the device word `out`, enqueued on `stream` (a hipStream_t, or null) with no allocation, no
   synchronisation and no memset. Its scratch is a fixed 512 KiB device array per GPU, shared by
   every call on that GPU: calls that overlap on one GPU race. Returns 0, or -1 with the reason in
   sf_last_error() (npts above 268435456, ncomp not 1 or 3, null pointers). */
int sf_norm_noalloc(const realw* field, long long npts, int ncomp, realw* out, void* stream);
void sf_batch_reset(long batch);
void sf_batch_time(long batch, int nsteps, double* stiffness_ms, double* step_ms);
void sf_batch_destroy(long batch);
/* host only: per-local-point flags of the fused batched mode for elements elems[q] of colour color[q]
   (colours run in increasing order) and the points Stacey or sources add to before the colours
   (forced[g] != 0): 1 = first writer, 2 = finisher of the Newmark update, 4 = finisher clears the
   acceleration; 125 entries per element of ibool0; returns -1 if two elements of a colour share a
   point */
int sf_batch_point_flags(int nglob, int nelem, const int* ibool0, const int* elems, const int* color,
                         const unsigned char* forced, unsigned char* flags);
/* launch order (0-based elastic elements, each once) of the atomic and coloured stiffness launches,
   e.g. a locality (Morton) order from the host; the ordered assembly keeps the processing list */
void sf_set_element_order(long* Mesh_pointer, int n, const int* order);
/* greedy first-fit element colouring of assembly = 2 (host only): color[q] for elems[q], elements
   sharing a global point differ; returns the number of colours or -1 (more than 32) */
int sf_color_elements(int nglob, int nelem, const int* ibool0, const int* elems, int* color);
/* jacobianstore (5,5,5,NSPEC_IRREGULAR) of the CPU solver for irregular elements (otherwise the
   inverse of det(d xi/d x) is recomputed from the nine derivatives) */
void sf_set_jacobian(long* Mesh_pointer, const realw* jacobianstore);
/* run any deferred corrector and wait for the mesh stream */
void sf_synchronize(long* Mesh_pointer);
/* HIP-event timing on the mesh stream of each elastic stiffness call (sf_set_tuning "timing_scope"
   0), or of its C-PML launch alone ("timing_scope" 1), for the next max_launches calls */
void sf_timing_start(long* Mesh_pointer, int max_launches);
void sf_timing_read(long* Mesh_pointer, double* total_ms, int* count);
/* message of the last fatal error */
const char* sf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SPECFEM3D_AMD_H */
