// pdp3 [ --version | -f path/to/parameters.xml ], src/main.cpp:6-80
#include <cstring>
#include <iostream>

#include "load_init_param.hpp"

namespace {
const char *option_value(int argc, char **argv, const char *opt) {
  for (int i = 1; i < argc; i++)
    if (std::strcmp(argv[i], opt) == 0) return i + 1 < argc ? argv[i + 1] : nullptr;
  return nullptr;
}
bool has_option(int argc, char **argv, const char *opt) {
  for (int i = 1; i < argc; i++)
    if (std::strcmp(argv[i], opt) == 0) return true;
  return false;
}
}  // namespace

int main(int argc, char **argv) {
  if (has_option(argc, argv, "-h")) {
    std::cerr << "USAGE:" << std::endl << "  pdp3 [ --version | -f path/to/parameters.xml ]" << std::endl;
    return 1;
  }
  if (has_option(argc, argv, "--version")) {
    std::cerr << "pdp3 " << pdp3_version() << std::endl;
    return 0;
  }
  const char *filename = "parameters.xml";
  if (has_option(argc, argv, "-f")) {
    filename = option_value(argc, argv, "-f");
    if (!filename) {
      std::cerr << "ERROR: configuration path is not specified" << std::endl;
      return 1;
    }
  }
  std::cout << std::endl << "Wellcome to pdp3, version: " << pdp3_version() << std::endl << std::endl;
  try {
    std::cerr << "INFO! Reading configuration file ``" << filename << "``" << std::endl;
    pdp3::LoadInitParam init_param(filename);
    std::cerr << "INFO! Initialization complete" << std::endl;
    init_param.run();
  } catch (const pdp3::ConfigError &e) {
    std::cerr << e.what() << std::endl;
    return e.code;
  } catch (const pdp3::Error &e) {
    std::cerr << "ERROR! " << e.what() << std::endl;
    return 1;
  } catch (const std::exception &e) {
    std::cerr << "ERROR! " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
