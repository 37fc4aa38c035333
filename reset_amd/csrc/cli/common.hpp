// Shared command-line handling of the dataset tools: `<exe> <input_folder>
// [--alpha=F] [--bucket=N]`, unknown flags rejected with exit status 1
// (triangle-count/TC.cu:223-233, baseline/tc/main.cu:46-78).
#pragma once

#include <cstdio>
#include <string>

#include "../../../include/reset_amd.hpp"

struct CliArgs {
  std::string folder;
  float load_factor = 0.2f;
  int bucket_size = 5;
};

// Returns false (after printing why) when the arguments are malformed.
inline bool parse_cli(int argc, char** argv, CliArgs& a) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s <input_folder> [--alpha=F] [--bucket=N]\n", argv[0]);
    return false;
  }
  a.folder = argv[1];
  for (int i = 2; i < argc; i++) {
    const std::string arg = argv[i];
    try {
      if (arg.rfind("--alpha=", 0) == 0) {
        a.load_factor = std::stof(arg.substr(8));
      } else if (arg.rfind("--bucket=", 0) == 0) {
        a.bucket_size = std::stoi(arg.substr(9));
      } else {
        std::fprintf(stderr, "Unknown argument: %s\n", arg.c_str());
        return false;
      }
    } catch (const std::exception&) {
      std::fprintf(stderr, "Bad value in argument: %s\n", arg.c_str());
      return false;
    }
  }
  std::printf("load_factor:%.2f , bucket_size:%d\n", a.load_factor, a.bucket_size);
  return true;
}
