// stubs.hip -- entry points of the reference interface outside the single-partition hot path.
//
// MPI halo exchange entry points are no-ops at NPROC=1 (no interfaces). Features outside this
// library's scope stop with a clear message, as the reference's own stubs do
// (src/gpu/specfem3D_gpu_cuda_method_stubs.c). register_host_array pins nothing here: page
// registration of caller memory is not used on this platform; transfers stage through the
// library's own pinned buffers instead.
#include "kernels.h"

using namespace sf;

// ---- MPI assembly at NPROC = 1 (assemble_MPI_scalar_cuda.cu, assemble_MPI_vector_cuda.cu)
extern "C" void FC_FUNC_(transfer_boun_pot_from_device, X)(long*, realw*, const int*) {}
extern "C" void FC_FUNC_(transfer_asmbl_pot_to_device, X)(long*, realw*, const int*) {}
extern "C" void FC_FUNC_(transfer_boun_accel_from_device, X)(long*, realw*, const int*) {}
extern "C" void FC_FUNC_(transfer_boundary_from_device_a, X)(long*, const int*) {}
extern "C" void FC_FUNC_(transfer_boundary_to_device_a, X)(long*, realw*, const int*, const int*) {}
extern "C" void FC_FUNC_(transfer_asmbl_accel_to_device, X)(long*, realw*, const int*, const int*, const int*,
                                                           const int*, const int*) {}
extern "C" void FC_FUNC_(transfer_sync_accel_to_device, X)(long*, realw*, const int*, const int*, const int*,
                                                          const int*, const int*) {}
extern "C" void FC_FUNC_(sync_copy_from_device, X)(long*, int*, realw*) {}

// ---- memory reports (check_fields_cuda.cu)
extern "C" void FC_FUNC_(get_free_device_memory, X)(realw* free_mb, realw* used_mb, realw* total_mb) {
  size_t fr = 0, tot = 0;
  SF_CHECK(hipMemGetInfo(&fr, &tot));
  *free_mb = (realw)(fr / 1048576.0);
  *total_mb = (realw)(tot / 1048576.0);
  *used_mb = *total_mb - *free_mb;
}
extern "C" void FC_FUNC_(output_free_device_memory, X)(int* myrank) {
  size_t fr = 0, tot = 0;
  SF_CHECK(hipMemGetInfo(&fr, &tot));
  if (myrank && *myrank == 0)
    std::printf("  GPU usage: used = %.1f MB, free = %.1f MB, total = %.1f MB\n", (tot - fr) / 1048576.0,
                fr / 1048576.0, tot / 1048576.0);
}
extern "C" void FC_FUNC_(register_host_array, X)(int*, realw*) {}
extern "C" void FC_FUNC_(unregister_host_array, X)(realw*) {}
extern "C" void FC_FUNC_(pause_for_debug, X)() {}

// ---- out of scope: stop with a message (parameter lists as the reference declares them)
#define SF_UNSUPPORTED(NAME, WHAT, ...) \
  extern "C" void FC_FUNC_(NAME, X)(__VA_ARGS__) { fatal(#NAME, WHAT " is not supported by this library"); }
SF_UNSUPPORTED(add_source_main_rec_noise_cu, "noise tomography", long*, int*, int*, int*)
SF_UNSUPPORTED(check_error_vectors, "debug check routine", int*, realw*, realw*)
SF_UNSUPPORTED(check_max_norm_b_accel, "debug check routine", int*, realw*, int*)
SF_UNSUPPORTED(check_max_norm_b_accel_gpu, "debug check routine", int*, realw*, long*, int*)
SF_UNSUPPORTED(check_max_norm_b_displ, "debug check routine", int*, realw*, int*)
SF_UNSUPPORTED(check_max_norm_b_displ_gpu, "debug check routine", int*, realw*, long*, int*)
SF_UNSUPPORTED(check_max_norm_b_veloc_gpu, "debug check routine", int*, realw*, long*, int*)
SF_UNSUPPORTED(check_max_norm_displ, "debug check routine", int*, realw*, int*)
SF_UNSUPPORTED(check_max_norm_displ_gpu, "debug check routine", int*, realw*, long*, int*)
SF_UNSUPPORTED(check_max_norm_vector, "debug check routine", int*, realw*, int*)
SF_UNSUPPORTED(compute_coupling_ocean_cuda, "ocean load", long*, int*)
SF_UNSUPPORTED(compute_kernels_hess_cuda, "Hessian kernels", long*, realw*, int*, int*)
SF_UNSUPPORTED(compute_kernels_strgth_noise_cu, "noise tomography", long*, realw*, realw*)
SF_UNSUPPORTED(compute_smooth_gpu, "kernel smoothing", long*, realw*, realw*, realw*, realw*, realw*, int*)
SF_UNSUPPORTED(fault_solver_gpu, "fault simulation", long*, long*, realw*, int*)
SF_UNSUPPORTED(get_max_accel, "debug check routine", int*, int*, long*)
SF_UNSUPPORTED(get_smooth_gpu, "kernel smoothing", long*, realw*)
SF_UNSUPPORTED(initialize_fault_data_gpu, "fault simulation", long*, int*, int*, int*, int*, int*)
SF_UNSUPPORTED(initialize_fault_solver_gpu, "fault simulation", long*, int*, realw*, realw*, int*)
SF_UNSUPPORTED(make_displ_rand, "noise tomography", long*, realw*)
SF_UNSUPPORTED(noise_read_add_surface_movie_cu, "noise tomography", long*, realw*, int*)
SF_UNSUPPORTED(prepare_fault_device, "fault simulation", long*, int*, realw*)
SF_UNSUPPORTED(prepare_fields_gravity_device, "gravity", long*, int*, realw*, realw*, realw*, realw*)
SF_UNSUPPORTED(prepare_fields_noise_device, "noise tomography", long*, int*, int*, int*, int*, int*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*)
SF_UNSUPPORTED(prepare_smooth_gpu, "kernel smoothing", long*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, int*, int*)
SF_UNSUPPORTED(transfer_compute_kernel_answers_from_device, "debug check routine", long*, realw*, int*, realw*, int*, realw*, int*)
SF_UNSUPPORTED(transfer_compute_kernel_fields_from_device, "debug check routine", long*, realw*, int*, realw*, int*, realw*, realw*, realw*, realw*, realw*, int*, realw*, realw*, realw*, realw*, realw*, int*, realw*, int*, realw*, int*, realw*, int*, realw*, realw*, int*)
SF_UNSUPPORTED(transfer_datat_to_host, "fault simulation", long*, int*, realw*, int*)
SF_UNSUPPORTED(transfer_fault_data_to_device, "fault simulation", long*, int*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, int*, int*, int*)
SF_UNSUPPORTED(transfer_fault_data_to_host, "fault simulation", long*, int*, int*, int*, realw*, realw*, realw*)
SF_UNSUPPORTED(transfer_kernels_hess_ac_tohost, "Hessian kernels", long*, realw*, realw*, realw*, int*)
SF_UNSUPPORTED(transfer_kernels_hess_el_tohost, "Hessian kernels", long*, realw*, realw*, realw*, realw*, int*)
SF_UNSUPPORTED(transfer_kernels_noise_to_host, "noise tomography", long*, realw*, int*)
SF_UNSUPPORTED(transfer_rsf_data_todevice, "fault simulation", long*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, int*)
SF_UNSUPPORTED(transfer_rsf_data_tohost, "fault simulation", long*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*, realw*)
SF_UNSUPPORTED(transfer_sigma_from_device, "noise tomography", int*, realw*, long*)
SF_UNSUPPORTED(transfer_surface_to_host, "noise tomography", long*, realw*)
SF_UNSUPPORTED(transfer_swf_data_todevice, "fault simulation", long*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*)
SF_UNSUPPORTED(transfer_swf_data_tohost, "fault simulation", long*, int*, int*, realw*, realw*, realw*, realw*, realw*, realw*)
