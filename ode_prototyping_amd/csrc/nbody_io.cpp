// Trajectory files in the reference's binary format (output_to_file,
// src/demo/nbody_io.hpp:4-22): u64 n_times, u64 n_var, f64 softening, then
// per saved state f64 t and n_var f64 values, native (little-endian) order.
#include <cstdint>
#include <cstdio>

#include "odehip.h"

extern "C" int odehip_write_trajectory(const char* path, size_t n_times,
                                       size_t n_var, double softening,
                                       const double* times,
                                       const double* states) {
  if (!path || (n_times && (!times || !states))) return ODEHIP_ERR_INVALID;
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return ODEHIP_ERR_INVALID;
  const uint64_t header[2] = {n_times, n_var};
  bool ok = std::fwrite(header, sizeof(uint64_t), 2, f) == 2 &&
            std::fwrite(&softening, sizeof(double), 1, f) == 1;
  for (size_t i = 0; ok && i < n_times; ++i)
    ok = std::fwrite(times + i, sizeof(double), 1, f) == 1 &&
         std::fwrite(states + i * n_var, sizeof(double), n_var, f) == n_var;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? ODEHIP_OK : ODEHIP_ERR_INVALID;
}
