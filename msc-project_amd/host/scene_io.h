// scene_io.h -- host-side scene input and image output of the reference's file formats.
//
// The reference reads a two-document YAML scene through yaml-cpp (Pathtracer.cpp:32-275), OBJ
// meshes through a modified tinyobjloader (PolygonObject.cpp:5-43) and writes OpenEXR half
// RGBA (main.cpp:197-208). None of those libraries is available here, so this module restates
// the parts the reference uses: a YAML subset that keeps repeated keys in document order,
// tinyobj's single-shape OBJ semantics, and an uncompressed scanline EXR writer.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace swr {
namespace io {

// ---- YAML subset ----------------------------------------------------------------------------
// Block and flow mappings and sequences, plain and quoted scalars, comments and '---'
// document separators. Mapping keys keep their order and may repeat, as yaml-cpp's iteration
// over the scene document does (Pathtracer.cpp:166-272).
struct Node {
  enum Kind { kNull, kScalar, kSeq, kMap } kind = kNull;
  std::string scalar;
  std::vector<Node> seq;
  std::vector<std::pair<std::string, Node>> map;

  bool is_map() const { return kind == kMap; }
  bool is_seq() const { return kind == kSeq; }
  size_t size() const { return kind == kMap ? map.size() : (kind == kSeq ? seq.size() : 0); }
  const Node* find(const std::string& key) const;  // first value of key, or nullptr
};

// Parses every document of a YAML text; throws std::runtime_error with a line number.
std::vector<Node> parse_yaml(const std::string& text);

// ---- OBJ ------------------------------------------------------------------------------------
// tinyobjloader as modified by the reference (tiny_obj_loader.cpp:643-890): one shape from the
// whole file; faces fan-triangulated; vertices deduplicated on their (v, vn, vt) triple and
// numbered in order of first use; negative indices relative; g/o/usemtl/mtllib ignored.
struct ObjMesh {
  std::vector<float> positions;  // xyz
  std::vector<float> normals;    // xyz, only for corners that carry a normal index
  std::vector<float> texcoords;  // uv, only for corners that carry a texture index
  std::vector<uint32_t> indices;
};
ObjMesh load_obj(const std::string& path);

// ---- EXR ------------------------------------------------------------------------------------
// IEEE binary16 with round-to-nearest-even (overflow to infinity, NaN kept).
uint16_t float_to_half(float f);
float half_to_float(uint16_t h);
// Single-part scanline OpenEXR, no compression, channels A, B, G, R as HALF, A = 1.
void write_exr_rgba(const std::string& path, const float* rgb, int width, int height);

// ---- texture images -------------------------------------------------------------------------
// StandardTexture's file (StandardTexture.h convert<>: the scene directory + "/" + string),
// decoded as OpenImageIO's texture() sees it: RGB floats, row 0 the top scanline (t = 0).
// Formats: PFM ("PF" RGB / "Pf" grey, float32, scanlines stored bottom-up, the scale's sign
// giving the byte order), binary PPM ("P6", 8- or 16-bit samples divided by maxval) and
// single-part scanline OpenEXR without compression (HALF, FLOAT or UINT channels, the data
// window as the image). Channels the file lacks read as 0, OpenImageIO's default fill: a
// one-channel image (PFM "Pf", EXR "Y") gives (v, 0, 0).
struct Image {
  int width = 0, height = 0;
  std::vector<float> rgb;
};
Image read_image(const std::string& path);

}  // namespace io
}  // namespace swr
