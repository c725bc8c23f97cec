// main.cpp -- openabl-hip, the .abl compiler with the -b hip backend.
//
// Command line, exit codes and the build/run/timing contract are the
// reference driver's (src/main.cpp:143-321, src/Cli.cpp:31-95), so the
// reference's harness can call it in place of OpenABL:
//   openabl-hip -i model.abl -o outdir -b hip [-P name=value] [-C key=value]
//               [-A assetdir] [-B] [-R] [--lint-only]
// Exit status 0 on success, 1 on an error, 2 when the backend does not
// support a feature the model uses.  With -R the generated build.sh and
// run.sh are executed and "Execution time: <seconds>s" (the wall time of
// run.sh) is printed, which bench/openabl.py of the reference parses.
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

#include "analysis.hpp"
#include "hip_backend.hpp"
#include "parser.hpp"

#ifndef ABL_DEFAULT_ASSET_DIR
#define ABL_DEFAULT_ASSET_DIR "./asset"
#endif

namespace {

struct Options {
  bool help = false, lint_only = false, build = false, run = false;
  std::string backend, input, output_dir, asset_dir, deps_dir;
  std::map<std::string, std::string> params, config;
};

struct OptionError : std::runtime_error {
  explicit OptionError(const std::string &m) : std::runtime_error(m) {}
};

std::pair<std::string, std::string> pair_of(const std::string &s, const char *what) {
  const size_t eq = s.find('=');
  if (eq == std::string::npos) throw OptionError(std::string("Malformed ") + what + ": Missing \"=\"");
  return {s.substr(0, eq), s.substr(eq + 1)};
}

Options parse_options(int argc, char **argv) {
  Options o;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      o.help = true;
      return o;
    }
    if (a == "--lint-only") {
      o.lint_only = true;
      continue;
    }
    if (a == "-B" || a == "--build") {
      o.build = true;
      continue;
    }
    if (a == "-R" || a == "--run") {
      o.run = true;
      continue;
    }
    if (i + 1 == argc) throw OptionError("Missing argument for option \"" + a + "\"");
    const std::string v = argv[++i];
    if (a == "-b" || a == "--backend") o.backend = v;
    else if (a == "-i" || a == "--input") o.input = v;
    else if (a == "-o" || a == "--output-dir") o.output_dir = v;
    else if (a == "-A" || a == "--asset-dir") o.asset_dir = v;
    else if (a == "-D" || a == "--deps-dir") o.deps_dir = v;
    else if (a == "-P" || a == "--param") o.params.insert(pair_of(v, "parameter"));
    else if (a == "-C" || a == "--config") o.config.insert(pair_of(v, "configuration value"));
    else throw OptionError("Unknown option \"" + a + "\"");
  }
  if (o.input.empty()) throw OptionError("Missing input file (-i or --input)");
  if (o.asset_dir.empty()) o.asset_dir = ABL_DEFAULT_ASSET_DIR;
  if (o.deps_dir.empty()) o.deps_dir = "./deps";
  if (o.lint_only) return o;
  if (o.backend.empty()) throw OptionError("Missing backend (-b or --backend)");
  return o;
}

void help() {
  std::cout << "Usage: ./openabl-hip -i input.abl -o ./output-dir -b hip\n\n"
               "Options:\n"
               "  -A, --asset-dir    Asset directory (default: the installed openabl_amd/asset)\n"
               "  -b, --backend      Backend\n"
               "  -B, --build        Build the generated code\n"
               "  -C, --config       Specify a configuration value (name=value)\n"
               "  -D, --deps         Deps directory (default: ./deps)\n"
               "  -h, --help         Display this help\n"
               "  -i, --input        Input file\n"
               "  -o, --output-dir   Output directory\n"
               "  -P, --param        Specify a simulation parameter (name=value)\n"
               "  -R, --run          Build and run the generated code\n"
               "\n"
               "Available backends:\n"
               " * hip (MI355X / gfx950)\n"
               "\n"
               "Available configuration options:\n"
               " * bool use_float (default: false)\n"
               " * string hip.program (default: the input file's base name)\n"
               " * hip.* runtime keys (include/abl_hip.h), applied by the generated program\n"
            << std::flush;
}

bool is_dir(const std::string &p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}
bool is_file(const std::string &p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

bool read_all(const std::string &path, std::string *out) {
  std::ifstream f(path);
  if (!f) return false;
  std::ostringstream s;
  s << f.rdbuf();
  *out = s.str();
  return true;
}

void make_dirs(const std::string &p) {
  std::string cur;
  std::stringstream ss(p);
  std::string part;
  if (!p.empty() && p[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    ::mkdir(cur.c_str(), 0755);
  }
}

std::string absolute(const std::string &p) {
  if (!p.empty() && p[0] == '/') return p;
  char buf[4096];
  if (!::getcwd(buf, sizeof buf)) return p;
  return std::string(buf) + "/" + p;
}

std::string base_name(const std::string &path) {
  std::string b = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
  const size_t dot = b.rfind('.');
  if (dot != std::string::npos && dot > 0) b = b.substr(0, dot);
  return b;
}

}  // namespace

int main(int argc, char **argv) {
  Options opt;
  try {
    opt = parse_options(argc, argv);
  } catch (const OptionError &e) {
    help();
    std::cerr << "\nERROR: " << e.what() << std::endl;
    return 1;
  }
  if (opt.help) {
    help();
    return 0;
  }
  std::string main_src, lib_src;
  if (!read_all(opt.input, &main_src)) {
    std::cerr << "File \"" << opt.input << "\" could not be opened." << std::endl;
    return 1;
  }
  if (!is_dir(opt.asset_dir)) {
    std::cerr << "Asset directory \"" << opt.asset_dir << "\" does not exist "
              << "(override with -A or --asset-dir)" << std::endl;
    return 1;
  }
  const std::string lib_path = opt.asset_dir + "/lib.abl";
  if (!read_all(lib_path, &lib_src)) {
    std::cerr << "Library file \"" << lib_path << "\" could not be opened." << std::endl;
    return 1;
  }
  std::string perr;
  auto script = ablc::parse(main_src, &perr);
  if (!script) {
    std::cerr << perr << std::endl;
    return 1;
  }
  auto lib = ablc::parse(lib_src, &perr);
  if (!lib) {
    std::cerr << perr << std::endl;
    return 1;
  }
  ablc::Diagnostics diag;
  diag.out = &std::cerr;
  ablc::Analysis analysis(*script, opt.params, diag, opt.backend);
  analysis.library(*lib);
  analysis.main_script();
  if (diag.count > 0) return 1;
  if (opt.lint_only) return 0;

  if (!opt.output_dir.empty()) {
    make_dirs(opt.output_dir);
  } else if (opt.build || opt.run) {
    char tmpl[] = "/tmp/openabl-XXXXXX";
    const char *d = ::mkdtemp(tmpl);
    if (!d) {
      std::cerr << "Could not create a temporary directory" << std::endl;
      return 1;
    }
    opt.output_dir = d;
    std::cout << "Writing to directory " << opt.output_dir << std::endl;
  } else {
    std::cerr << "Missing output directory (-o or --output-dir)" << std::endl;
    return 1;
  }
  opt.deps_dir = absolute(opt.deps_dir);
  if (opt.backend != "hip") {
    std::cerr << "Unknown backend \"" << opt.backend << "\"" << std::endl;
    return 1;
  }
  ablc::BackendContext ctx;
  ctx.output_dir = opt.output_dir;
  ctx.asset_dir = absolute(opt.asset_dir);
  ctx.deps_dir = opt.deps_dir;
  ctx.config = opt.config;
  ctx.input_name = base_name(opt.input);
  try {
    ablc::HipBackend().generate(*script, analysis, ctx);
  } catch (const ablc::BackendError &e) {
    std::cerr << e.what() << std::endl;
    return 2;
  } catch (const std::runtime_error &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  if (opt.build || opt.run) {
    if (::chdir(opt.output_dir.c_str()) != 0) {
      std::cerr << "Could not enter " << opt.output_dir << std::endl;
      return 1;
    }
    if (!is_file("./build.sh")) {
      std::cerr << "Build file for this backend not found" << std::endl;
      return 1;
    }
    if (std::system("./build.sh") != 0) {
      std::cerr << "Build failed" << std::endl;
      return 1;
    }
    if (opt.run) {
      if (!is_file("./run.sh")) {
        std::cerr << "Run file for this backend not found" << std::endl;
        return 1;
      }
      const auto start = std::chrono::high_resolution_clock::now();
      if (std::system("./run.sh") != 0) {
        std::cerr << "Run failed" << std::endl;
        return 1;
      }
      const auto end = std::chrono::high_resolution_clock::now();
      const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count();
      std::cout << "Execution time: " << ms / 1000.0 << "s" << std::endl;
    }
  }
  return 0;
}
