// io_capi.cpp - C-ABI entry points of the scenario readers (pbf.h).
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/pbf.h"
#include "io/ConfigReader.hpp"
#include "io/PartReader.hpp"

namespace pbf {
void set_error_text(const char *text);  // pbf_api.cpp
}
using pbf::set_error_text;

extern "C" int pbf_read_par(const char *path, pbf_params *out,
                            char *part_file, size_t cap) {
  if (!path || !out) {
    set_error_text("NULL argument");
    return PBF_ERR_INVALID;
  }
  try {
    ConfigReader reader;
    const ConfigParameters p = reader.read(path);
    *out = p.toPbfParams();
    if (part_file && cap > 0) {
      std::strncpy(part_file, p.partInputFile.c_str(), cap - 1);
      part_file[cap - 1] = '\0';
    }
  } catch (const std::exception &e) {
    set_error_text(e.what());
    return PBF_ERR_INVALID;
  }
  set_error_text("");
  return PBF_OK;
}

extern "C" int pbf_read_particles(const char *path, pbf_particle *out,
                                  uint32_t cap, uint32_t *n) {
  if (!path || !n) {
    set_error_text("NULL argument");
    return PBF_ERR_INVALID;
  }
  try {
    PartReader reader;
    const std::vector<Particle> parts = reader.read(path);
    *n = (uint32_t)parts.size();
    if (out) {
      const size_t k = std::min<size_t>(cap, parts.size());
      std::memcpy(out, parts.data(), k * sizeof(pbf_particle));
    }
  } catch (const std::exception &e) {
    set_error_text(e.what());
    return PBF_ERR_INVALID;
  }
  set_error_text("");
  return PBF_OK;
}
