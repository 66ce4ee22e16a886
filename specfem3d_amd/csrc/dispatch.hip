// dispatch.hip -- GLL constants, exact/fast kernel dispatch, error reporting.
#include <cstring>
#include <string>

#include "kernels.h"

namespace sf {

GLLk g_gll_host;

void set_gll_constants(const realw* hprime, const realw* hprimewgll, const realw* wgllwgll) {
  std::memcpy(g_gll_host.hp, hprime, sizeof(g_gll_host.hp));
  std::memcpy(g_gll_host.hpw, hprimewgll, sizeof(g_gll_host.hpw));
  std::memcpy(g_gll_host.wxy, wgllwgll, sizeof(g_gll_host.wxy));
}

// variants compiled from k_elastic.hip / k_acoustic.hip with and without FMA contraction
void launch_elastic_exact(hipStream_t, const ElasticArgs&, bool, bool, int, int);
void launch_elastic_fast(hipStream_t, const ElasticArgs&, bool, bool, int, int);
void launch_element_strain_exact(hipStream_t, const ElasticArgs&, realw*, realw*, bool);
void launch_element_strain_fast(hipStream_t, const ElasticArgs&, realw*, realw*, bool);
void launch_elastic_batch_exact(hipStream_t, const BatchArgs&, int, int, int);
void launch_elastic_batch_fast(hipStream_t, const BatchArgs&, int, int, int);
void launch_elastic_pml_exact(hipStream_t, const PmlArgs&, int);
void launch_elastic_pml_fast(hipStream_t, const PmlArgs&, int);
void launch_acoustic_exact(hipStream_t, const AcousticArgs&, int);
void launch_acoustic_fast(hipStream_t, const AcousticArgs&, int);

void launch_elastic(hipStream_t s, const ElasticArgs& a, bool strain, bool att, bool exact, int mode,
                    int att_variant) {
  if (exact)
    launch_elastic_exact(s, a, strain, att, mode, att_variant);
  else
    launch_elastic_fast(s, a, strain, att, mode, att_variant);
}

void launch_element_strain(hipStream_t s, const ElasticArgs& a, realw* eps, realw* eps_tr3, bool with_trace,
                           bool exact) {
  if (exact)
    launch_element_strain_exact(s, a, eps, eps_tr3, with_trace);
  else
    launch_element_strain_fast(s, a, eps, eps_tr3, with_trace);
}

void launch_elastic_pml(hipStream_t s, const PmlArgs& a, bool exact, int mode) {
  if (exact)
    launch_elastic_pml_exact(s, a, mode);
  else
    launch_elastic_pml_fast(s, a, mode);
}

void launch_visco_batch_exact(hipStream_t, const BatchArgs&, int, int);
void launch_visco_batch_fast(hipStream_t, const BatchArgs&, int, int);
void launch_visco_batch(hipStream_t s, const BatchArgs& a, int sc, int mode, bool exact) {
  if (exact)
    launch_visco_batch_exact(s, a, sc, mode);
  else
    launch_visco_batch_fast(s, a, sc, mode);
}

void launch_elastic_batch(hipStream_t s, const BatchArgs& a, int sc, int mode, int overlay, bool exact) {
  if (exact)
    launch_elastic_batch_exact(s, a, sc, mode, overlay);
  else
    launch_elastic_batch_fast(s, a, sc, mode, overlay);
}

void launch_acoustic_batch_exact(hipStream_t, const BatchAcArgs&, int, int);
void launch_acoustic_batch_fast(hipStream_t, const BatchAcArgs&, int, int);
void launch_acoustic_batch(hipStream_t s, const BatchAcArgs& a, int sc, int mode, bool exact) {
  if (exact)
    launch_acoustic_batch_exact(s, a, sc, mode);
  else
    launch_acoustic_batch_fast(s, a, sc, mode);
}

void launch_acoustic(hipStream_t s, const AcousticArgs& a, bool exact, int mode) {
  if (exact)
    launch_acoustic_exact(s, a, mode);
  else
    launch_acoustic_fast(s, a, mode);
}

static std::string g_last_error;

void set_error(const char* msg) { g_last_error = msg ? msg : ""; }

[[noreturn]] void fatal(const char* where, const char* msg) {
  char buf[1024];
  std::snprintf(buf, sizeof(buf), "specfem3d_amd error: %s: %s", where, msg);
  g_last_error = buf;
  std::fprintf(stderr, "%s\n", buf);
  // like the reference's exit_on_error: append to OUTPUT_FILES/error_message*.txt if present
  if (FILE* f = std::fopen("OUTPUT_FILES/error_message000000.txt", "a")) {
    std::fprintf(f, "%s\n", buf);
    std::fclose(f);
  }
  std::fflush(stderr);
  std::exit(EXIT_FAILURE);
}

}  // namespace sf

extern "C" const char* sf_last_error(void) { return sf::g_last_error.c_str(); }
