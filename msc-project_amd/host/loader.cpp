// loader.cpp -- decodes the reference's scene files into a device scene + render description
// (swr_scene_load_yaml), and writes images (swr_write_exr).
//
// Follows Pathtracer::construct (src/core/Pathtracer.cpp:32-275) and the convert<> decoders:
// document 0 holds image/settings/camera/filter/sampler with the reference's defaults, document 1
// is one mapping whose repeated keys (object, shader, light) are read in order. Decoders are as
// strict as the reference's: a block with the wrong number of keys is an error.
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

#include "scene_io.h"
#include "swr.h"
#include "swr_internal.h"

namespace swr {
namespace {

using io::Node;

[[noreturn]] void bad(const std::string& what) { throw std::runtime_error(what); }

const Node& need(const Node& n, const char* key, const char* ctx) {
  const Node* v = n.find(key);
  if (!v) bad(std::string(ctx) + ": missing key '" + key + "'");
  return *v;
}

void need_map(const Node& n, size_t keys, const char* ctx) {
  if (!n.is_map() || n.size() != keys)
    bad(std::string(ctx) + ": expected a map of " + std::to_string(keys) + " keys, got " + std::to_string(n.size()));
}

std::string as_str(const Node& n) {
  if (n.kind != Node::kScalar) bad("expected a scalar");
  return n.scalar;
}

int as_int(const Node& n) {
  const std::string s = as_str(n);
  size_t used = 0;
  const long v = std::stol(s, &used, 0);
  if (used != s.size()) bad("'" + s + "' is not an integer");
  return (int)v;
}

float as_float(const Node& n) {
  const std::string s = as_str(n);
  size_t used = 0;
  const float v = std::stof(s, &used);
  if (used != s.size()) bad("'" + s + "' is not a number");
  return v;
}

void as_vec(const Node& n, float* out, size_t count, const char* ctx) {
  if (!n.is_seq() || n.size() != count)
    bad(std::string(ctx) + ": expected a sequence of " + std::to_string(count) + " numbers");
  for (size_t k = 0; k < count; ++k) out[k] = as_float(n.seq[k]);
}

const float kPi180 = 0.01745329251994329577f;

void rot(int axis, float deg, float m[9]) {
  const float a = deg * kPi180, c = std::cos(a), s = std::sin(a);
  std::memset(m, 0, 9 * sizeof(float));
  const int i = (axis + 1) % 3, j = (axis + 2) % 3;
  m[axis * 3 + axis] = 1.f;
  m[i * 3 + i] = c;
  m[i * 3 + j] = -s;
  m[j * 3 + i] = s;
  m[j * 3 + j] = c;
}

void mul(const float a[9], const float b[9], float o[9]) {
  float r[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  std::memcpy(o, r, sizeof r);
}

std::string dir_of(const std::string& path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

void decode_setup(const Node& setup, swr_render_desc* d) {
  // Defaults (Pathtracer.cpp:45-161, ThinLensCamera.h:22-28).
  std::memset(d, 0, sizeof *d);
  d->width = 700;
  d->height = 500;
  d->base = 8;
  d->min_depth = 2;
  d->max_depth = 100;
  d->threshold = 0.01f;
  d->shading_size = 4096;
  d->bin_exponent = 25;
  d->sampler = SWR_SAMPLER_STRATIFIED;
  d->filter = SWR_FILTER_TENT;
  d->camera.type = SWR_CAMERA_THINLENS;
  d->camera.focal_length = 55.f;
  d->camera.aperture = 5.6f;
  d->camera.focal_distance = 1000.f;
  if (!setup.is_map()) return;
  if (const Node* n = setup.find("image")) {  // Image.h:37-50
    need_map(*n, 3, "image");
    d->width = as_int(need(*n, "width", "image"));
    d->height = as_int(need(*n, "height", "image"));
    d->base = as_int(need(*n, "sample base", "image"));
  }
  if (const Node* n = setup.find("settings")) {  // Settings.h:22-37
    need_map(*n, 6, "settings");
    d->min_depth = as_int(need(*n, "min depth", "settings"));
    d->max_depth = as_int(need(*n, "max depth", "settings"));
    d->threshold = as_float(need(*n, "threshold", "settings"));
    (void)as_int(need(*n, "bucket size", "settings"));
    d->shading_size = as_int(need(*n, "shading size", "settings"));
    d->bin_exponent = as_int(need(*n, "bin exponent", "settings"));
  }
  if (const Node* n = setup.find("camera")) {
    const std::string type = as_str(need(*n, "type", "camera"));
    if (type == "ThinLens") {  // ThinLensCamera.h convert<>
      need_map(*n, 6, "ThinLens camera");
      d->camera.type = SWR_CAMERA_THINLENS;
      as_vec(need(*n, "translation", "camera"), d->camera.translation, 3, "camera translation");
      as_vec(need(*n, "rotation", "camera"), d->camera.rotation, 3, "camera rotation");
      d->camera.focal_length = as_float(need(*n, "focal length", "camera"));
      d->camera.focal_distance = as_float(need(*n, "focal distance", "camera"));
      d->camera.aperture = as_float(need(*n, "aperture", "camera"));
    } else if (type == "PinHole") {  // PinHoleCamera.h convert<>
      need_map(*n, 4, "PinHole camera");
      d->camera.type = SWR_CAMERA_PINHOLE;
      as_vec(need(*n, "translation", "camera"), d->camera.translation, 3, "camera translation");
      as_vec(need(*n, "rotation", "camera"), d->camera.rotation, 3, "camera rotation");
      d->camera.focal_length = as_float(need(*n, "focal length", "camera"));
    }
  }
  if (const Node* n = setup.find("filter")) {
    const std::string type = as_str(need(*n, "type", "filter"));
    if (type == "Tent" || type == "Box") {
      need_map(*n, 1, "filter");
      d->filter = type == "Tent" ? SWR_FILTER_TENT : SWR_FILTER_BOX;
    }
  }
  if (const Node* n = setup.find("sampler")) {
    const std::string type = as_str(need(*n, "type", "sampler"));
    if (type == "Stratified" || type == "Independent" || type == "Grid") {
      need_map(*n, 1, "sampler");
      d->sampler = type == "Stratified" ? SWR_SAMPLER_STRATIFIED
                                        : (type == "Independent" ? SWR_SAMPLER_INDEPENDENT : SWR_SAMPLER_GRID);
    }
  }
}

// Constant texture colour (ConstantTexture.h convert<>).
void decode_texture(const Node& t, float colour[3]) {
  const std::string type = as_str(need(t, "type", "texture"));
  if (type != "Constant") bad("texture type '" + type + "' is only read as a shader's texture");
  if (t.size() != 2) bad("Constant texture: expected 2 keys");
  as_vec(need(t, "colour", "texture"), colour, 3, "texture colour");
}

void check(swr_context* ctx, int rc) {
  if (rc < 0) throw Error{rc, ctx->err};
}

// A texture graph as scene texture nodes (convert<> of StandardTexture.h, LayeredTexture.h and
// ConstantTexture.h); returns the root node. Standard textures name an image file relative to
// the scene directory; each file is decoded once.
struct TextureLoader {
  swr_context* ctx;
  swr_scene* scene;
  std::string dir;
  std::map<std::string, int> images;

  int node(const Node& t) {
    const std::string type = as_str(need(t, "type", "texture"));
    swr_texture_desc d{};
    if (type == "Constant") {
      if (t.size() != 2) bad("Constant texture: expected 2 keys");
      d.type = SWR_TEXTURE_CONSTANT;
      as_vec(need(t, "colour", "texture"), d.colour, 3, "texture colour");
    } else if (type == "Standard") {
      if (t.size() != 2) bad("Standard texture: expected 2 keys");
      const std::string file = dir + "/" + as_str(need(t, "string", "texture"));
      auto it = images.find(file);
      if (it == images.end()) {
        io::Image im;
        try {
          im = io::read_image(file);
        } catch (const std::exception& e) {
          throw Error{SWR_E_IO, e.what()};
        }
        const int idx = swr_scene_add_image(scene, im.width, im.height, im.rgb.data());
        check(ctx, idx);
        it = images.emplace(file, idx).first;
      }
      d.type = SWR_TEXTURE_IMAGE;
      d.image = it->second;
    } else if (type == "Layered") {
      if (t.size() != 4) bad("Layered texture: expected 4 keys");
      d.type = SWR_TEXTURE_LAYERED;
      d.upper = node(need(t, "upper", "texture"));
      d.lower = node(need(t, "lower", "texture"));
      d.mask = node(need(t, "mask", "texture"));
    } else {
      bad("unknown texture type '" + type + "'");
    }
    const int idx = swr_scene_add_texture(scene, &d);
    check(ctx, idx);
    return idx;
  }
};

void load_scene(swr_context* ctx, const std::string& path, swr_scene** out, swr_render_desc* desc) {
  std::ifstream f(path);
  if (!f) throw Error{SWR_E_IO, "cannot open scene file " + path};
  std::stringstream ss;
  ss << f.rdbuf();
  std::vector<Node> docs;
  try {
    docs = io::parse_yaml(ss.str());
  } catch (const std::exception& e) {
    throw Error{SWR_E_PARSE, e.what()};
  }
  if (docs.size() < 2) throw Error{SWR_E_PARSE, "scene file needs a setup and a scene document"};
  swr_scene* scene = nullptr;
  check(ctx, swr_scene_create(ctx, &scene));
  try {
    try {
      decode_setup(docs[0], desc);
    } catch (const std::exception& e) {
      throw Error{SWR_E_PARSE, std::string("setup: ") + e.what()};
    }
    const Node& root = docs[1];
    if (!root.is_map()) throw Error{SWR_E_PARSE, "scene document must be a mapping"};
    const std::string dir = dir_of(path);
    TextureLoader textures{ctx, scene, dir, {}};
    int num_shaders = 0;
    // Pass 1: objects and shaders in document order (Pathtracer.cpp:166-217).
    for (const auto& kv : root.map) {
      const Node& v = kv.second;
      try {
        if (kv.first == "object" && as_str(need(v, "type", "object")) == "Polygon") {
          need_map(v, 6, "Polygon object");  // PolygonObject.h convert<>
          const std::string file = dir + "/" + as_str(need(v, "filename", "object"));
          float t[3], r[3], s[3];
          as_vec(need(v, "translation", "object"), t, 3, "object translation");
          as_vec(need(v, "rotation", "object"), r, 3, "object rotation");
          as_vec(need(v, "scale", "object"), s, 3, "object scale");
          const int shader = as_int(need(v, "shader", "object"));
          io::ObjMesh m;
          try {
            m = io::load_obj(file);
          } catch (const std::exception& e) {
            throw Error{SWR_E_IO, e.what()};
          }
          // T * Rx * Ry * Rz * S (PolygonObject.cpp:17-42)
          float rx[9], ry[9], rz[9], tmp[9], R[9];
          rot(0, r[0], rx);
          rot(1, r[1], ry);
          rot(2, r[2], rz);
          mul(rx, ry, tmp);
          mul(tmp, rz, R);
          const size_t nv = m.positions.size() / 3;
          std::vector<float> pos4(4 * nv);
          for (size_t i = 0; i < nv; ++i) {
            const float p[3] = {m.positions[3 * i] * s[0], m.positions[3 * i + 1] * s[1], m.positions[3 * i + 2] * s[2]};
            for (int k = 0; k < 3; ++k) pos4[4 * i + k] = R[3 * k] * p[0] + R[3 * k + 1] * p[1] + R[3 * k + 2] * p[2] + t[k];
            pos4[4 * i + 3] = 0.f;
          }
          const bool has_uv = m.texcoords.size() == 2 * nv && nv > 0;
          check(ctx, swr_scene_add_mesh(scene, pos4.data(), (uint32_t)nv, m.indices.data(),
                                        (uint32_t)(m.indices.size() / 3), has_uv ? m.texcoords.data() : nullptr,
                                        shader));
        } else if (kv.first == "shader") {
          const std::string type = as_str(need(v, "type", "shader"));
          swr_shader_desc sd{};
          if (type == "Lambert") {  // LambertShader.h convert<>
            need_map(v, 3, "Lambert shader");
            sd.type = SWR_SHADER_LAMBERT;
          } else if (type == "Glossy") {  // GGX extension: Lambert's keys plus roughness
            need_map(v, 4, "Glossy shader");
            sd.type = SWR_SHADER_GLOSSY;
            sd.roughness = as_float(need(v, "roughness", "shader"));
          } else {
            continue;  // unknown shader types are skipped, as in the reference
          }
          // A Constant texture is the shader's colour; other textures become texture nodes.
          const Node& tex = need(v, "texture", "shader");
          const bool constant = as_str(need(tex, "type", "texture")) == "Constant";
          if (constant) decode_texture(tex, sd.colour);
          sd.reflectance = as_float(need(v, "reflectance", "shader"));
          check(ctx, swr_scene_add_shader(scene, &sd));
          if (!constant) check(ctx, swr_scene_set_shader_texture(scene, num_shaders, textures.node(tex)));
          ++num_shaders;
        } else if (kv.first == "environment") {  // constant environment extension
          float rgb[3];
          decode_texture(v, rgb);
          check(ctx, swr_scene_set_environment(scene, rgb));
        }
      } catch (const Error&) {
        throw;
      } catch (const std::exception& e) {
        throw Error{SWR_E_PARSE, kv.first + ": " + e.what()};
      }
    }
    // Pass 2: lights (Pathtracer.cpp:219-272).
    for (const auto& kv : root.map) {
      if (kv.first != "light") continue;
      const Node& v = kv.second;
      try {
        if (as_str(need(v, "type", "light")) != "Quad") continue;
        need_map(v, 5, "Quad light");  // QuadLight.h convert<>
        swr_quad_light_desc ld{};
        as_vec(need(v, "translation", "light"), ld.translation, 3, "light translation");
        as_vec(need(v, "rotation", "light"), ld.rotation, 3, "light rotation");
        as_vec(need(v, "scale", "light"), ld.scale, 2, "light scale");
        ld.intensity = as_float(need(v, "intensity", "light"));
        check(ctx, swr_scene_add_quad_light(scene, &ld));
      } catch (const Error&) {
        throw;
      } catch (const std::exception& e) {
        throw Error{SWR_E_PARSE, std::string("light: ") + e.what()};
      }
    }
    check(ctx, swr_scene_commit(scene));
  } catch (...) {
    swr_scene_destroy(scene);
    throw;
  }
  *out = scene;
}

}  // namespace
}  // namespace swr

extern "C" {

int swr_scene_load_yaml(swr_context* ctx, const char* path, swr_scene** out_scene, swr_render_desc* out_desc) {
  if (!ctx || !path || !out_scene || !out_desc) return SWR_E_INVALID;
  *out_scene = nullptr;
  try {
    swr::load_scene(ctx, path, out_scene, out_desc);
    return SWR_OK;
  } catch (const swr::Error& e) {
    ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return SWR_E_PARSE;
  }
}

int swr_write_exr(const char* path, const float* rgb, int32_t width, int32_t height) {
  if (!path || !rgb || width <= 0 || height <= 0) return SWR_E_INVALID;
  try {
    swr::io::write_exr_rgba(path, rgb, width, height);
    return SWR_OK;
  } catch (const std::exception&) {
    return SWR_E_IO;
  }
}

}  // extern "C"
