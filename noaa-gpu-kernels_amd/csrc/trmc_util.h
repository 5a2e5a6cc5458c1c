// Internal host helpers shared by the library's translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <string>

#include "trmc.h"

namespace trmc {

// thread-local last error (trmc_last_error)
void set_error(const std::string &msg);
int fail(int code, const char *fmt, ...);

// generator, formats, statistics (trmc_gen.cpp)
int generate_network(int kind, int n_segments, int n_reaches, uint64_t seed, trmc_topo *out);
int generate_params(int n, uint64_t seed, trmc_chparams_f64 *out);
int generate_forcing(int n, int T, int M, uint64_t seed, double *qlat);
int generate_forcing_members(int n, int T, int m0, int M, uint64_t seed, double *qlat);
int generate_state(int n, const trmc_chparams_f64 *p, double q0, double *qdp0, double *dp0);
int load_topo(const char *path, trmc_topo *t, std::string &err);
int save_topo(const char *path, const trmc_topo *t);
int compute_stats(const double *pred, const double *truth, int n, double *out);
int concat_topo(const trmc_topo *nets, int n, trmc_topo *out, int *seg_offset, int *reach_offset);

}  // namespace trmc
