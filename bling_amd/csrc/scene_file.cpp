// Flat binary scene files: a bgpu_scene_desc written as one file, so a
// scene flattened once (by the .bling reader, a generator or the Haskell
// front end) is loaded again without parsing.
//
// Layout (little-endian, every section starting on a 64-byte boundary):
//   FileHeader (512 B)
//   vertices  f32[3 nv]     normals f32[3 nv] (flag)   uvs f32[2 nv] (flag)
//   indices   i32[3 nt]     tri_mesh i32[nt] (flag)
//   meshes    bgpu_mesh[num_meshes]
//   shapes    bgpu_shape[num_shapes]
//   materials bgpu_material[num_materials]
//   lights    bgpu_light[num_lights] (radiance pointers stored as 0)
//   one radiance map per light, f32[16 w h] (none for a sun/sky map)
//   tex_nodes bgpu_tex_node[num_tex_nodes]
//   images    bgpu_image[num_images] (texel pointers stored as 0)
//   the texels of each image, f32[channels w h]
// The header records the record sizes; a file written with other record
// layouts is refused rather than misread.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "bgpu_internal.h"

namespace bgpu {
namespace {

constexpr char kMagic[8] = {'B', 'G', 'P', 'U', 'S', 'C', 'N', '1'};
constexpr uint32_t kVersion = 5;  // 2: three spectra and four scalars per material; 3: texture nodes, images;
                                  // 4: point and directional lights (bgpu_light::vec);
                                  // 5: sun/sky maps (bgpu_light::map, ::sky)
constexpr uint32_t kByteOrder = 0x01020304u;
constexpr uint32_t kHasNormals = 1, kHasUvs = 2, kHasTriMesh = 4;
constexpr uint64_t kAlign = 64;
constexpr uint32_t kMaxCount = 0x7fffffffu;

struct FileHeader {
  char magic[8];
  uint32_t version, byte_order;
  uint32_t num_vertices, num_triangles, num_meshes, num_shapes, num_materials, num_lights;
  uint32_t flags, pad0;
  uint32_t size_mesh, size_shape, size_material, size_light, size_camera, size_film;
  uint64_t file_bytes;
  bgpu_camera camera;
  bgpu_film film;
  uint32_t num_tex_nodes, num_images, size_tex_node, size_image;
};
static_assert(sizeof(FileHeader) <= 512, "header must fit its 512 bytes");
constexpr uint64_t kHeaderBytes = 512;

uint64_t up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Byte offsets of every section, from the header's counts.  Returns false
// when a light map's size is out of range.
struct Sections {
  uint64_t vertices, normals, uvs, indices, tri_mesh, meshes, shapes, materials, lights;
  std::vector<uint64_t> maps;
  uint64_t tex_nodes, images;
  std::vector<uint64_t> texels;
  uint64_t end;
};

// A light evaluated from its parameters stores no radiance map.
bool has_map(const bgpu_light& l) { return !(l.type == BGPU_LIGHT_INFINITE && l.map == BGPU_LIGHTMAP_SUNSKY); }

// Texels of an image record, or 0 when its size is out of range.
uint64_t image_floats(const bgpu_image& im) {
  if (im.width <= 0 || im.height <= 0 || (im.channels != 1 && im.channels != 16)) return 0;
  if ((uint64_t)im.width * (uint64_t)im.height > (1ull << 28)) return 0;
  return (uint64_t)im.channels * im.width * im.height;
}

// Offsets of every section; light maps and texel blocks only once their
// records are known (`lights`, `images` non-null).  Returns false when a map
// or an image is out of range.
bool layout(const FileHeader& h, const std::vector<bgpu_light>* lights, const std::vector<bgpu_image>* images,
            Sections* s) {
  uint64_t o = kHeaderBytes;
  auto take = [&](uint64_t bytes) {
    uint64_t at = o;
    o = up(o + bytes);
    return at;
  };
  const uint64_t nv = h.num_vertices, nt = h.num_triangles;
  s->vertices = take(12 * nv);
  s->normals = (h.flags & kHasNormals) ? take(12 * nv) : 0;
  s->uvs = (h.flags & kHasUvs) ? take(8 * nv) : 0;
  s->indices = take(12 * nt);
  s->tri_mesh = (h.flags & kHasTriMesh) ? take(4 * nt) : 0;
  s->meshes = take(sizeof(bgpu_mesh) * (uint64_t)h.num_meshes);
  s->shapes = take(sizeof(bgpu_shape) * (uint64_t)h.num_shapes);
  s->materials = take(sizeof(bgpu_material) * (uint64_t)h.num_materials);
  s->lights = take(sizeof(bgpu_light) * (uint64_t)h.num_lights);
  s->maps.clear();
  s->texels.clear();
  s->tex_nodes = s->images = s->end = 0;
  if (!lights) return true;
  for (const bgpu_light& l : *lights) {
    if (!has_map(l)) {
      s->maps.push_back(0);
      continue;
    }
    if (l.width <= 0 || l.height <= 0 || (uint64_t)l.width * (uint64_t)l.height > (1ull << 28)) return false;
    s->maps.push_back(take(64ull * l.width * l.height));
  }
  s->tex_nodes = take(sizeof(bgpu_tex_node) * (uint64_t)h.num_tex_nodes);
  s->images = take(sizeof(bgpu_image) * (uint64_t)h.num_images);
  if (!images) return true;
  for (const bgpu_image& im : *images) {
    const uint64_t n = image_floats(im);
    if (!n) return false;
    s->texels.push_back(take(4 * n));
  }
  s->end = o;
  return true;
}

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

// A descriptor read from a file and the bytes it points into, in one block.
struct OwnedDesc {
  bgpu_scene_desc desc;  // first member: its address is the block's
  std::vector<unsigned char> bytes;
  std::vector<bgpu_light> lights;
  std::vector<bgpu_image> images;
};
static_assert(std::is_standard_layout<OwnedDesc>::value, "desc must sit at the block's address");

}  // namespace
}  // namespace bgpu

using namespace bgpu;

extern "C" int bgpu_scene_write(const bgpu_scene_desc* d, const char* path) {
  if (!d || !path) return fail(BGPU_E_INVALID, "bgpu_scene_write: NULL argument");
  if (d->num_vertices > kMaxCount || d->num_triangles > kMaxCount)
    return fail(BGPU_E_INVALID, "bgpu_scene_write: too many vertices or triangles");
  if ((d->num_vertices && !d->vertices) || (d->num_triangles && !d->indices) ||
      (d->num_meshes && !d->meshes) || (d->num_shapes && !d->shapes) ||
      (d->num_materials && !d->materials) || (d->num_lights && !d->lights) ||
      (d->num_tex_nodes && !d->tex_nodes) || (d->num_images && !d->images))
    return fail(BGPU_E_INVALID, "bgpu_scene_write: a counted array is NULL");
  if (d->num_tex_nodes > kMaxCount || d->num_images > (1u << 16))
    return fail(BGPU_E_INVALID, "bgpu_scene_write: too many texture nodes or images");
  FileHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.byte_order = kByteOrder;
  h.num_vertices = d->num_vertices;
  h.num_triangles = d->num_triangles;
  h.num_meshes = d->num_meshes;
  h.num_shapes = d->num_shapes;
  h.num_materials = d->num_materials;
  h.num_lights = d->num_lights;
  h.flags = (d->normals ? kHasNormals : 0) | (d->uvs ? kHasUvs : 0) | (d->tri_mesh ? kHasTriMesh : 0);
  h.size_mesh = sizeof(bgpu_mesh);
  h.size_shape = sizeof(bgpu_shape);
  h.size_material = sizeof(bgpu_material);
  h.size_light = sizeof(bgpu_light);
  h.size_camera = sizeof(bgpu_camera);
  h.size_film = sizeof(bgpu_film);
  h.camera = d->camera;
  h.film = d->film;
  h.num_tex_nodes = d->num_tex_nodes;
  h.num_images = d->num_images;
  h.size_tex_node = sizeof(bgpu_tex_node);
  h.size_image = sizeof(bgpu_image);
  std::vector<bgpu_light> lights(d->lights, d->lights + d->num_lights);
  for (const bgpu_light& l : lights)
    if (has_map(l) && !l.radiance) return fail(BGPU_E_INVALID, "bgpu_scene_write: light without a radiance map");
  std::vector<bgpu_image> images(d->images, d->images + d->num_images);
  for (const bgpu_image& im : images)
    if (!im.texels) return fail(BGPU_E_INVALID, "bgpu_scene_write: image without texels");
  Sections s;
  if (!layout(h, &lights, &images, &s))
    return fail(BGPU_E_INVALID, "bgpu_scene_write: light map or image size out of range");
  h.file_bytes = s.end;

  File f(std::fopen(path, "wb"));
  if (!f) return fail(BGPU_E_INVALID, std::string("bgpu_scene_write: cannot open ") + path);
  uint64_t at = 0;
  bool ok = true;
  auto put = [&](uint64_t off, const void* p, uint64_t bytes) {
    static const unsigned char zeros[kAlign] = {};
    while (ok && at < off) {  // padding up to the section
      uint64_t k = std::min<uint64_t>(off - at, kAlign);
      ok = std::fwrite(zeros, 1, k, f.get()) == k;
      at += k;
    }
    if (ok && bytes) ok = std::fwrite(p, 1, bytes, f.get()) == bytes;
    at += bytes;
  };
  unsigned char head[kHeaderBytes] = {};
  std::memcpy(head, &h, sizeof(h));
  put(0, head, kHeaderBytes);
  const uint64_t nv = d->num_vertices, nt = d->num_triangles;
  put(s.vertices, d->vertices, 12 * nv);
  if (d->normals) put(s.normals, d->normals, 12 * nv);
  if (d->uvs) put(s.uvs, d->uvs, 8 * nv);
  put(s.indices, d->indices, 12 * nt);
  if (d->tri_mesh) put(s.tri_mesh, d->tri_mesh, 4 * nt);
  put(s.meshes, d->meshes, sizeof(bgpu_mesh) * (uint64_t)d->num_meshes);
  put(s.shapes, d->shapes, sizeof(bgpu_shape) * (uint64_t)d->num_shapes);
  put(s.materials, d->materials, sizeof(bgpu_material) * (uint64_t)d->num_materials);
  std::vector<bgpu_light> stored = lights;
  for (bgpu_light& l : stored) l.radiance = nullptr;
  put(s.lights, stored.data(), sizeof(bgpu_light) * stored.size());
  for (size_t i = 0; i < lights.size(); ++i)
    if (has_map(lights[i])) put(s.maps[i], lights[i].radiance, 64ull * lights[i].width * lights[i].height);
  put(s.tex_nodes, d->tex_nodes, sizeof(bgpu_tex_node) * (uint64_t)d->num_tex_nodes);
  std::vector<bgpu_image> stored_images = images;
  for (bgpu_image& im : stored_images) im.texels = nullptr;
  put(s.images, stored_images.data(), sizeof(bgpu_image) * stored_images.size());
  for (size_t i = 0; i < images.size(); ++i) put(s.texels[i], images[i].texels, 4 * image_floats(images[i]));
  put(s.end, nullptr, 0);
  if (!ok || std::fflush(f.get()) != 0) return fail(BGPU_E_INVALID, std::string("bgpu_scene_write: write failed: ") + path);
  return BGPU_OK;
}

extern "C" int bgpu_scene_read(const char* path, bgpu_scene_desc** out) {
  if (!path || !out) return fail(BGPU_E_INVALID, "bgpu_scene_read: NULL argument");
  *out = nullptr;
  File f(std::fopen(path, "rb"));
  if (!f) return fail(BGPU_E_INVALID, std::string("bgpu_scene_read: cannot open ") + path);
  if (std::fseek(f.get(), 0, SEEK_END) != 0) return fail(BGPU_E_INVALID, "bgpu_scene_read: cannot seek");
  long size = std::ftell(f.get());
  std::rewind(f.get());
  if (size < (long)kHeaderBytes) return fail(BGPU_E_INVALID, "bgpu_scene_read: file shorter than its header");
  std::unique_ptr<OwnedDesc> od(new (std::nothrow) OwnedDesc());
  if (!od) return fail(BGPU_E_NOMEM, "bgpu_scene_read: out of host memory");
  try {
    od->bytes.resize((size_t)size);
  } catch (...) {
    return fail(BGPU_E_NOMEM, "bgpu_scene_read: out of host memory");
  }
  if (std::fread(od->bytes.data(), 1, (size_t)size, f.get()) != (size_t)size)
    return fail(BGPU_E_INVALID, "bgpu_scene_read: read failed");
  FileHeader h;
  std::memcpy(&h, od->bytes.data(), sizeof(h));
  if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(BGPU_E_INVALID, "bgpu_scene_read: not a scene file");
  if (h.version != kVersion || h.byte_order != kByteOrder)
    return fail(BGPU_E_INVALID, "bgpu_scene_read: unsupported version or byte order");
  if (h.size_mesh != sizeof(bgpu_mesh) || h.size_shape != sizeof(bgpu_shape) ||
      h.size_material != sizeof(bgpu_material) || h.size_light != sizeof(bgpu_light) ||
      h.size_camera != sizeof(bgpu_camera) || h.size_film != sizeof(bgpu_film) ||
      h.size_tex_node != sizeof(bgpu_tex_node) || h.size_image != sizeof(bgpu_image))
    return fail(BGPU_E_INVALID, "bgpu_scene_read: record layouts differ from this library's");
  if (h.num_vertices > kMaxCount || h.num_triangles > kMaxCount || h.num_meshes > kMaxCount ||
      h.num_shapes > kMaxCount || h.num_materials > kMaxCount || h.num_lights > (1u << 16) ||
      h.num_tex_nodes > kMaxCount || h.num_images > (1u << 16) ||
      (h.flags & ~(kHasNormals | kHasUvs | kHasTriMesh)))
    return fail(BGPU_E_INVALID, "bgpu_scene_read: corrupt header");
  if (h.file_bytes != (uint64_t)size) return fail(BGPU_E_INVALID, "bgpu_scene_read: file size differs from its header");
  // Light records first, then image records: the sections after each depend
  // on them.
  Sections s;
  if (!layout(h, nullptr, nullptr, &s) || s.lights + sizeof(bgpu_light) * (uint64_t)h.num_lights > (uint64_t)size)
    return fail(BGPU_E_INVALID, "bgpu_scene_read: file shorter than its sections");
  od->lights.resize(h.num_lights);
  if (h.num_lights) std::memcpy(od->lights.data(), od->bytes.data() + s.lights, sizeof(bgpu_light) * h.num_lights);
  if (!layout(h, &od->lights, nullptr, &s)) return fail(BGPU_E_INVALID, "bgpu_scene_read: light map size out of range");
  if (s.images + sizeof(bgpu_image) * (uint64_t)h.num_images > (uint64_t)size)
    return fail(BGPU_E_INVALID, "bgpu_scene_read: file shorter than its sections");
  od->images.resize(h.num_images);
  if (h.num_images) std::memcpy(od->images.data(), od->bytes.data() + s.images, sizeof(bgpu_image) * h.num_images);
  if (!layout(h, &od->lights, &od->images, &s)) return fail(BGPU_E_INVALID, "bgpu_scene_read: image size out of range");
  if (s.end != (uint64_t)size) return fail(BGPU_E_INVALID, "bgpu_scene_read: file size differs from its sections");
  unsigned char* b = od->bytes.data();
  bgpu_scene_desc& d = od->desc;
  std::memset(&d, 0, sizeof(d));
  d.num_vertices = h.num_vertices;
  d.num_triangles = h.num_triangles;
  d.vertices = reinterpret_cast<const float*>(b + s.vertices);
  d.normals = (h.flags & kHasNormals) ? reinterpret_cast<const float*>(b + s.normals) : nullptr;
  d.uvs = (h.flags & kHasUvs) ? reinterpret_cast<const float*>(b + s.uvs) : nullptr;
  d.indices = reinterpret_cast<const int32_t*>(b + s.indices);
  d.tri_mesh = (h.flags & kHasTriMesh) ? reinterpret_cast<const int32_t*>(b + s.tri_mesh) : nullptr;
  d.num_meshes = h.num_meshes;
  d.meshes = reinterpret_cast<const bgpu_mesh*>(b + s.meshes);
  d.num_shapes = h.num_shapes;
  d.shapes = reinterpret_cast<const bgpu_shape*>(b + s.shapes);
  d.num_materials = h.num_materials;
  d.materials = reinterpret_cast<const bgpu_material*>(b + s.materials);
  for (size_t i = 0; i < od->lights.size(); ++i)
    od->lights[i].radiance = has_map(od->lights[i]) ? reinterpret_cast<const float*>(b + s.maps[i]) : nullptr;
  d.num_lights = h.num_lights;
  d.lights = od->lights.data();
  d.camera = h.camera;
  d.film = h.film;
  d.num_tex_nodes = h.num_tex_nodes;
  d.tex_nodes = reinterpret_cast<const bgpu_tex_node*>(b + s.tex_nodes);
  for (size_t i = 0; i < od->images.size(); ++i) od->images[i].texels = reinterpret_cast<const float*>(b + s.texels[i]);
  d.num_images = h.num_images;
  d.images = od->images.data();
  *out = &od.release()->desc;
  return BGPU_OK;
}

extern "C" int bgpu_scene_desc_free(bgpu_scene_desc* desc) {
  if (!desc) return fail(BGPU_E_INVALID, "bgpu_scene_desc_free: NULL descriptor");
  delete reinterpret_cast<OwnedDesc*>(desc);
  return BGPU_OK;
}

extern "C" int bgpu_scene_load(bgpu_ctx* ctx, const char* path, bgpu_scene** out) {
  if (!ctx || !path || !out) return fail(BGPU_E_INVALID, "bgpu_scene_load: NULL argument");
  *out = nullptr;
  bgpu_scene_desc* d = nullptr;
  int rc = bgpu_scene_read(path, &d);
  if (rc) return rc;
  rc = bgpu_scene_create(ctx, d, out);
  std::string err = rc ? bgpu_last_error() : std::string();
  bgpu_scene_desc_free(d);
  if (rc) return fail(rc, err);
  return BGPU_OK;
}
