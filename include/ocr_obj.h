/*
 * ocr_obj.h — scene ingest for the draw path: the reference's Wavefront OBJ +
 * MTL loader (obj_load, obj_load.cpp:181-568) and the PNG decoding its
 * textures need (texture::set_location + sf::Image::loadFromFile,
 * texture.cpp:133-136, :642).  Host code in libocr_engine.so; the C++ facade's
 * objects_container uses it as its default load function, as the reference's
 * constructor does (objects_container.cpp:22).
 *
 * The loader keeps the reference's parse rules, quirks included:
 *   - attribute lines (v, vt, vn) are read with atoi when the line holds no
 *     '.', with atof otherwise, each field from the character after the
 *     previous space (obj_load.cpp:144-165);
 *   - faces must be triangles with v/vt/vn indices (obj_load.cpp:168-186);
 *     indices outside the lists are an error here (undefined behaviour in
 *     the reference);
 *   - every usemtl starts a part that runs to the next usemtl; faces before
 *     the first usemtl belong to no part (obj_load.cpp:417-425);
 *   - a part's texture is the first map_Kd after its newmtl (matched by
 *     name prefix, not stopping at the next newmtl), else its Kd colour as a
 *     32x32 texture, else magenta (obj_load.cpp:427-500);
 *   - a part's bump map is the map_Bump inside its own newmtl block.
 */
#ifndef OCR_OBJ_H_INCLUDED
#define OCR_OBJ_H_INCLUDED

#include <stdint.h>

#include "ocr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OCR_OBJ_TEX_FILE    0   /* texture_path names an existing file      */
#define OCR_OBJ_TEX_COLOUR  1   /* 32x32 of colour (Kd * 255, truncated)    */
#define OCR_OBJ_TEX_MAGENTA 2   /* no usable map_Kd or Kd                   */

/* One usemtl group: an object of the container (obj_load.cpp:503-548). */
typedef struct ocr_obj_part {
    uint32_t first_tri;         /* into the loaded triangle array          */
    uint32_t ntri;
    int32_t  texture_kind;      /* OCR_OBJ_TEX_*                            */
    uint8_t  colour[4];         /* RGBA of a colour / magenta texture       */
    char     texture_cache[512];/* the texture cache key (texture_context::make_new_cached) */
    char     texture_path[512]; /* the file to decode (OCR_OBJ_TEX_FILE)    */
    char     bump_path[512];    /* "" when the material has no map_Bump     */
    char     group_name[512];   /* "g" name when usemtl and g lines pair up */
} ocr_obj_part;
/* The strings of a part are NUL-terminated and cut at 511 bytes (the
 * reference keeps std::strings). */

typedef struct ocr_obj_mesh ocr_obj_mesh;

/* Loads `path` (and the .mtl beside it) with vertex positions scaled by
 * `scale` (objects_container::requested_scale).  OCR_ERR_INVALID if the
 * file cannot be opened or a face index is out of range. */
int      ocr_obj_load(const char* path, float scale, ocr_obj_mesh** out);
uint32_t ocr_obj_ntri(const ocr_obj_mesh* m);
uint32_t ocr_obj_nparts(const ocr_obj_mesh* m);
/* Copies the triangles (object_id 0, vertex_col 0; the draw path sets the
 * object id when the context is built) and the parts. */
int      ocr_obj_get(const ocr_obj_mesh* m, ocr_triangle* tris, ocr_obj_part* parts);
void     ocr_obj_free(ocr_obj_mesh* m);

/* Decodes a non-interlaced 8-bit PNG (grey, grey+alpha, RGB, RGBA or
 * palette with tRNS) into tightly packed RGBA8 rows, top row first (the
 * order sf::Image gives).  *rgba is released with ocr_png_free. */
int      ocr_png_load(const char* path, uint32_t* width, uint32_t* height, uint8_t** rgba);
void     ocr_png_free(uint8_t* rgba);

#ifdef __cplusplus
}
#endif
#endif
