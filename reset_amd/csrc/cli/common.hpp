// Shared command-line handling of the dataset tools: `<exe> <input_folder>
// [--alpha=F] [--bucket=N] [--method=a,b,...]`, unknown flags and methods
// rejected with exit status 1 (triangle-count/TC.cu:223-233,
// baseline/tc/main.cu:46-78, baseline/sss/main.cu:36-86).
#pragma once

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../../include/reset_amd.hpp"

struct CliArgs {
  std::string folder;
  float load_factor = 0.2f;
  int bucket_size = 5;
  std::vector<std::string> methods{"original"};
  bool has(const char* m) const { return std::find(methods.begin(), methods.end(), m) != methods.end(); }
};

// Returns false (after printing why) when the arguments are malformed.
// `allowed` lists the tool's methods; "all" selects every one of them.
inline bool parse_cli(int argc, char** argv, CliArgs& a,
                      const std::vector<std::string>& allowed = {"original"}) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s <input_folder> [--alpha=F] [--bucket=N] [--method=...]\n", argv[0]);
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
      } else if (arg.rfind("--method=", 0) == 0) {
        a.methods.clear();
        std::string list = arg.substr(9);
        size_t pos = 0;
        while (pos <= list.size()) {
          const size_t comma = list.find(',', pos);
          const std::string m = list.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
          if (m == "all") {
            a.methods = allowed;
          } else if (!m.empty()) {
            if (std::find(allowed.begin(), allowed.end(), m) == allowed.end()) {
              std::string all;
              for (const auto& x : allowed) all += (all.empty() ? "" : ", ") + x;
              std::fprintf(stderr, "Unknown method: %s (available: %s)\n", m.c_str(), all.c_str());
              return false;
            }
            if (!a.has(m.c_str())) a.methods.push_back(m);
          }
          if (comma == std::string::npos) break;
          pos = comma + 1;
        }
        if (a.methods.empty()) a.methods = {"original"};
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
