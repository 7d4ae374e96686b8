/*
 * raymond-amd host-side helpers above the C ABI (include/rm.h).
 *
 * The reference keeps material code generation in its Swift host
 * (host/scene/MaterialBuilder.swift:32-904); this is the C++ restatement a
 * host binds next to rm.h.  Input: the Blender exporter's node-graph JSON
 * (scripts/blender_exporter/raymond_blender/materials.py:300-560) wrapped as
 *   {"surface": [{"name": ..., "nodes": {...}}, ...],
 *    "light":   [{"name": ..., "nodes": {...}}, ...]}
 * where array position is the material index.  Output: source text defining
 * rm::shade_surface and rm::shade_light, to pass to rm_scene_compile_materials.
 */
#ifndef RAYMOND_AMD_RM_HOST_H
#define RAYMOND_AMD_RM_HOST_H

#include <stddef.h>
#include <stdint.h>
#include "rm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Returns 0 and a malloc'd source string, or -1 and a malloc'd error message
 * (cycle, multi-input link, unsupported node; MaterialBuilder.swift:45-51).
 * noinline != 0 emits material functions as out-of-line calls. */
int rmh_codegen(const char* materials_json, int noinline, char** out_source, char** out_error);
/* The texture slots the generated code reads, as a malloc'd JSON array in
 * slot order (the reference's registerTexture, MaterialBuilder.swift:
 * 329-435): {"slot", "kind": "image" with "path" (the exporter's file path),
 * or "nishita" with "sun_elevation", "altitude", "air_density",
 * "dust_density", "ozone_density" of a 512 x 512 sky texture to fill with
 * rmh_sky_nishita_texture}.  The caller passes the loaded or generated
 * images to rm_scene_create in that order.  Returns 0, or -1 and a malloc'd
 * error as rmh_codegen. */
int rmh_codegen_textures(const char* materials_json, char** out_json, char** out_error);
/* Identifier the codegen derives from a node, socket or material name
 * ("Base Color" -> "baseColor", MaterialBuilder.swift:19-43). */
int rmh_identifier(const char* name, char* out, size_t capacity);
void rmh_free(void* p);
/* Builds a BVH8 over n triangles (9 floats each) on the host and verifies it:
 * each triangle referenced once, every quantised child box containing its
 * subtree.  Returns 0, the node count and the depth; -1 and a message otherwise.
 * (A test hook for the builder behind rm_scene_build_accel.) */
int rmh_bvh8_check(const float* tris, uint32_t n, int threads, uint32_t* out_nodes, uint32_t* out_depth, char* err,
                   size_t err_cap);
/* Builds the two-level form of rm_scene_build_accel_two_level on the host
 * from scene arrays (vertices 3 floats each, faces 3 indices relative to the
 * instance's vertex_offset, instance descriptions) and traces n rays (8
 * floats: origin, tmin, direction, tmax) through it on the CPU with the
 * device traversal code: closest hits (any = 0) as rm_trace_closest writes
 * them, or the first hit found (any = 1).  out_stored receives the
 * triangles stored (each shape once).  Returns 0, or -1 and a message.  (A
 * test hook for the traversal behind the two-level form.) */
int rmh_two_level_trace(const float* vertices, uint32_t num_vertices, const uint32_t* indices, uint32_t num_faces,
                        const rm_instance_desc* instances, uint32_t num_instances, const float* rays, uint32_t n,
                        int any, int threads, float* t, uint32_t* inst, uint32_t* prim, float* bary,
                        uint64_t* out_stored, char* err, size_t err_cap);
/* Compiles generated material source for gfx950 with hiprtc, without a device,
 * into the code-object cache that rm_scene_compile_materials reads (the
 * directory in $RM_JIT_CACHE, else ~/.cache/raymond_amd).  Same source and
 * options give the same cache entry, so a later scene load skips the compile.
 * Returns 0 (out_hit = 1 if the entry already existed) or RM_ERROR_COMPILE
 * with the hiprtc log in `log`.  No reference counterpart: the reference
 * compiles its Metal library at scene load (SceneLoader.swift:62-75). */
int rmh_jit_precompile(const char* source, size_t length, const char* options, int* out_hit, char* log,
                       size_t log_capacity);

/* A mesh as the Blender exporter writes it (shapes.py:19-123): per vertex
 * position, normal and uv; per face three vertex indices and the slot of the
 * face's material in the shape's "materials" list. */
typedef struct rmh_ply_mesh {
    uint32_t num_vertices;
    uint32_t num_faces;
    float* vertices;        /* 3 per vertex */
    float* normals;         /* 3 per vertex */
    float* texcoords;       /* 2 per vertex */
    uint32_t* indices;      /* 3 per face */
    uint8_t* face_palette;  /* per face: index into the shape's material list */
    float bounds_min[3];    /* as the reference's reader computes them (see below) */
    float bounds_max[3];
} rmh_ply_mesh;

/* Reads an ASCII PLY file of the exporter's exact layout.  Replaces the
 * reference's PLYReader (raymond/io/PLYReader.m:1-169) as ShapeBuilder uses it
 * (host/scene/ShapeBuilder.swift:150-187, 260-283): the same header lines,
 * numbers parsed as (float)digits * 10^-decimals in float arithmetic, and
 * bounds whose maximum is only raised by a coordinate that is not a new
 * minimum.  Returns 0 with the arrays malloc'd into *out (free with
 * rmh_ply_free), or -1 and a malloc'd message (free with rmh_free) where the
 * reference would assert: a header line that differs, a face that is not a
 * triangle, a malformed number. */
int rmh_ply_read(const char* path, rmh_ply_mesh* out, char** out_error);
void rmh_ply_free(rmh_ply_mesh* mesh);

/* Writes an RGBA float image (width*height*4, row 0 first) as an OpenEXR
 * file: RGB times norm, alpha 1, as 32-bit float channels A, B, G, R,
 * uncompressed scanlines.  Replaces Renderer.saveFrame's EXR output
 * (host/Renderer.swift:804-807, utils/Metal.swift:75-124), which the caller
 * reproduces by passing the accumulation buffer and norm = 1 / (frames - 1)
 * when accumulating, else 1.  Returns 0, or -1 and a malloc'd message (free
 * with rmh_free). */
int rmh_save_exr(const char* path, const float* rgba, uint32_t width, uint32_t height, float norm,
                 char** out_error);

/* Nishita sky (replaces host/sky/sky_nishita.cpp as SkyLoader.mm:8-39 and
 * MaterialBuilder.swift:329-378 use it).  rmh_sky_nishita_texture fills a
 * width x height RGBA float texture (alpha untouched) with the CIE XYZ sky
 * radiance, texel (x, y) at azimuth 2 pi x / width - pi and elevation
 * (pi/2 + pi/(4 height)) (y / height)^2; rmh_sky_nishita_sun gives the XYZ
 * radiance of the sun disc's lower and upper edge.  Angles in radians,
 * altitude in metres.  Return 0, or -1 on a null or empty argument. */
int rmh_sky_nishita_texture(float* pixels, int width, int height, float sun_elevation, float altitude,
                            float air_density, float dust_density, float ozone_density);
int rmh_sky_nishita_sun(float sun_elevation, float angular_diameter, float altitude, float air_density,
                        float dust_density, float* bottom, float* top);

#ifdef __cplusplus
}
#endif
#endif
