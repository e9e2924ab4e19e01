// main.cpp -- command-line renderer: swr [-h] [-o output.exr] scene.yaml
// Same command line as the reference (src/main.cpp:18-81). With -o it renders one iteration and
// writes half RGBA EXR with A = 1 (main.cpp:197-208); without -o it renders one iteration and
// prints its statistics (the reference would open its progressive viewer instead).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>

#include "pathtracer.h"

static void usage(FILE* f) {
  std::fprintf(f, "usage: swr [options] scene_file.yaml\noptions:\n  -h [ --help ]         produce help message\n"
                  "  -o [ --output ] arg   output as file\n");
}

int main(int argc, char** argv) {
  std::string input, output;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return EXIT_SUCCESS;
    } else if (a == "-o" || a == "--output") {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "error: the option '--output' requires an argument\n\n");
        usage(stderr);
        return EXIT_FAILURE;
      }
      output = argv[++i];
    } else {
      input = a;
    }
  }
  if (input.empty()) {
    std::fprintf(stderr, "error: a scene file is required but missing\n\n");
    usage(stderr);
    return EXIT_FAILURE;
  }
  struct stat st;
  if (stat(input.c_str(), &st) != 0) {
    std::fprintf(stderr, "error: the scene file does not exist\n\n");
    usage(stderr);
    return EXIT_FAILURE;
  }
  try {
    swr::Pathtracer pt(input);
    float* pixels;
    int w, h;
    pt.image(&pixels, &w, &h);
    const int it = pt.process();
    const swr_stats& s = pt.stats();
    std::printf("iteration %d: %dx%d, %llu closest + %llu shadow rays in %.2f ms (%.1f Mrays/s device)\n", it, w, h,
                (unsigned long long)s.closest_rays, (unsigned long long)s.shadow_rays, s.ms_total,
                (double)(s.closest_rays + s.shadow_rays) / (s.ms_total * 1e3));
    if (!output.empty()) {
      if (swr_write_exr(output.c_str(), pixels, w, h) != SWR_OK) {
        std::fprintf(stderr, "error: cannot write %s\n", output.c_str());
        return EXIT_FAILURE;
      }
      std::printf("Output file named \"%s\" written.\n", output.c_str());
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
