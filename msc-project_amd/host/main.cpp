// main.cpp -- command-line renderer: swr [-h] [-o output.exr] [-n iterations] [-p preview.ppm] scene.yaml
// Same command line as the reference (src/main.cpp:18-81), plus -n and -p for running headless.
//
// With -o it renders one iteration and writes half RGBA EXR with A = 1 (main.cpp:197-208).
// Without -o it runs the progressive loop of the reference's viewer (main.cpp:104-124, 150-183)
// without a window: process() until terminated -- Ctrl-C stands in for the viewer's ESC key and
// window close (main.cpp:83-102) -- or until -n iterations are done. After each iteration, -p
// writes the image the viewer would show, min(1, sum / iteration) * 255 as 8-bit RGB
// (main.cpp:163-168), to a binary PPM.
#include <algorithm>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "pathtracer.h"

static volatile std::sig_atomic_t g_stop = 0;
static void on_interrupt(int) { g_stop = 1; }

static void usage(FILE* f) {
  std::fprintf(f,
               "usage: swr [options] scene_file.yaml\noptions:\n  -h [ --help ]         produce help message\n"
               "  -o [ --output ] arg   output as file\n"
               "  -n [ --iterations ] arg  without -o: stop after this many iterations (default: until Ctrl-C)\n"
               "  -p [ --preview ] arg  without -o: write the viewer's 8-bit image as PPM after each iteration\n");
}

// The viewer's display conversion (main.cpp:163-168): per channel min(1, sum / iteration) * 255,
// truncated to an unsigned char.
static bool write_preview(const std::string& path, const float* sum, int w, int h, int iteration) {
  std::vector<unsigned char> rgb(3 * (size_t)w * h);
  for (size_t i = 0; i < rgb.size(); ++i) {
    const float v = std::min(1.f, sum[i] / (float)iteration) * 255.f;
    rgb[i] = static_cast<unsigned char>(v);
  }
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "P6\n%d %d\n255\n", w, h);
  const size_t n = std::fwrite(rgb.data(), 1, rgb.size(), f);
  const int rc = std::fclose(f);
  if (n != rgb.size() || rc != 0) return false;
  return std::rename(tmp.c_str(), path.c_str()) == 0;  // a reader never sees a partial frame
}

static bool arg_value(int argc, char** argv, int* i, const char* name, std::string* out) {
  if (*i + 1 >= argc) {
    std::fprintf(stderr, "error: the option '--%s' requires an argument\n\n", name);
    usage(stderr);
    return false;
  }
  *out = argv[++*i];
  return true;
}

int main(int argc, char** argv) {
  std::string input, output, preview, iterations;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage(stdout);
      return EXIT_SUCCESS;
    } else if (a == "-o" || a == "--output") {
      if (!arg_value(argc, argv, &i, "output", &output)) return EXIT_FAILURE;
    } else if (a == "-p" || a == "--preview") {
      if (!arg_value(argc, argv, &i, "preview", &preview)) return EXIT_FAILURE;
    } else if (a == "-n" || a == "--iterations") {
      if (!arg_value(argc, argv, &i, "iterations", &iterations)) return EXIT_FAILURE;
    } else {
      input = a;
    }
  }
  if (input.empty()) {
    std::fprintf(stderr, "error: a scene file is required but missing\n\n");
    usage(stderr);
    return EXIT_FAILURE;
  }
  long max_iterations = 0;  // 0: until interrupted
  if (!iterations.empty()) {
    char* end = nullptr;
    max_iterations = std::strtol(iterations.c_str(), &end, 10);
    if (*end || max_iterations <= 0) {
      std::fprintf(stderr, "error: --iterations needs a positive integer\n\n");
      usage(stderr);
      return EXIT_FAILURE;
    }
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
    if (!output.empty()) {
      const int it = pt.process();
      const swr_stats& s = pt.stats();
      std::printf("iteration %d: %dx%d, %llu closest + %llu shadow rays in %.2f ms (%.1f Mrays/s device)\n", it, w,
                  h, (unsigned long long)s.closest_rays, (unsigned long long)s.shadow_rays, s.ms_total,
                  (double)(s.closest_rays + s.shadow_rays) / (s.ms_total * 1e3));
      if (swr_write_exr(output.c_str(), pixels, w, h) != SWR_OK) {
        std::fprintf(stderr, "error: cannot write %s\n", output.c_str());
        return EXIT_FAILURE;
      }
      std::printf("Output file named \"%s\" written.\n", output.c_str());
      return EXIT_SUCCESS;
    }
    std::signal(SIGINT, on_interrupt);
    std::signal(SIGTERM, on_interrupt);
    while (pt.active()) {
      const int it = pt.process();
      if (it == 0) break;
      const swr_stats& s = pt.stats();
      std::printf("Iteration %d is complete (%.2f ms, %.1f Mrays/s device).\n", it, s.ms_total,
                  (double)(s.closest_rays + s.shadow_rays) / (s.ms_total * 1e3));
      std::fflush(stdout);
      if (!preview.empty() && !write_preview(preview, pixels, w, h, it)) {
        std::fprintf(stderr, "error: cannot write %s\n", preview.c_str());
        return EXIT_FAILURE;
      }
      if (g_stop || (max_iterations && it >= max_iterations)) pt.terminate();
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
