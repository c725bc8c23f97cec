// hip_backend.hpp -- the -b hip backend (hip_backend.cpp).
#pragma once

#include <map>
#include <stdexcept>
#include <string>

#include "analysis.hpp"
#include "ast.hpp"

namespace ablc {

// A feature the backend does not support: the driver exits with status 2,
// as the reference does for its BackendError (src/main.cpp:270-274).
struct BackendError : std::runtime_error {
  explicit BackendError(const std::string &m) : std::runtime_error(m) {}
};

// What a backend is handed (src/backend/Backend.hpp:26-31 of the reference
// passes the same three directories and the -C configuration).
struct BackendContext {
  std::string output_dir;
  std::string asset_dir;
  std::string deps_dir;
  std::map<std::string, std::string> config;
  std::string input_name;  // base name of the .abl file (the program's default name)
};

struct HipBackend {
  // writes main.cpp, steps.hip, abl_gen.h, build.sh, run.sh into the output dir
  void generate(Script &script, const Analysis &analysis, const BackendContext &ctx);
};

}  // namespace ablc
