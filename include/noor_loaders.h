/*
 * noor_loaders.h -- payload producers for files: Wavefront OBJ + MTL models
 * and RGB(A) float images.  They feed noor_scene.h (meshes, materials,
 * lights) and noor_payload_view.textures; nothing here touches a GPU.
 *
 * Replaces the reference's file-driven host layer:
 *   ObjLoader::load / getVertices / getNormals / getUVs / getLights
 *                                   (noor/source/objLoader.cpp:28-147)
 *   ObjLoader material getters      (noor/header/objLoader.h:60-169)
 *   ImageTexture::load              (noor/source/texture.cpp:49-76)
 * The reference parses OBJ through the glm_obj library, which it does not
 * ship; the behaviour restated here is that library's published one
 * (glmReadOBJ, glmUnitize, glmFacetNormals, glmVertexNormals), see
 * noor_obj_load.  FBX (Autodesk SDK), progressive JPEG, BMP, PSD and GIF
 * (stb_image), and the EXR compressions beyond ZIP are not provided.
 */
#ifndef NOOR_LOADERS_H
#define NOOR_LOADERS_H

#include "noor_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NOOR_PATH_MAX 256

/* One MTL material as the reference's ObjLoader getters report it
 * (noor/header/objLoader.h:90-169). */
typedef struct noor_obj_material {
    char name[64];
    uint32_t type;                 /* NOOR_MAT_*: illum 1 glossy, 3 mirror, 5 emitter, 6 metal,
                                      7 glass, anything else diffuse */
    float diffuse[3];              /* Kd */
    float specular[3];             /* Ks */
    float transmission[3];         /* Tf */
    float emittance[3];            /* Ke */
    float ior;                     /* Ni */
    float roughness;               /* 1 / Ns */
    float transparency;            /* d (or 1 - Tr) */
    float bumpfactor;              /* -bm of map_bump / bump, else 1 */
    float uvscale[2];              /* -s of map_Kd, else 1 1 */
    char map_kd[NOOR_PATH_MAX], map_ks[NOOR_PATH_MAX], map_ke[NOOR_PATH_MAX];
    char map_d[NOOR_PATH_MAX], map_bump[NOOR_PATH_MAX];   /* relative to the MTL file, "" if none */
} noor_obj_material;

/* The loaded model: one mesh with unified vertices (one per distinct
 * position / normal / texture-coordinate triple of a triangle corner). */
typedef struct noor_obj_view {
    uint32_t num_vertices;
    const float* positions;        /* xyz */
    const float* normals;          /* xyz */
    const float* uvs;              /* uv; (0, 0) where a corner has none */
    uint32_t num_triangles;
    const uint32_t* indices;       /* 3 per triangle, in file order (polygons as fans) */
    const uint32_t* material;      /* per triangle; num_materials (a default diffuse material) if none */
    uint32_t num_materials;
    const noor_obj_material* materials;
    uint32_t num_lights;           /* groups whose name contains "light" */
    const float* lights;           /* 15 floats each: base, u, v, n, Ke (a quad light) */
    float bbox_min[3], bbox_max[3];/* of the file's positions, before unitizing */
    float unitize_scale;           /* 1 if not unitized */
} noor_obj_view;

typedef struct noor_obj noor_obj;

/* Reads `path` (and the MTL files it names).  Polygons are split into fans
 * (v0, vi, vi+1); negative indices count back from the last element.  With
 * `unitize`, positions are centred on the box centre and scaled by
 * 2 / max(|max_x| + |min_x|, |max_y| + |min_y|, |max_z| + |min_z|) (glmUnitize).
 * Normals are recomputed (glmFacetNormals, glmVertexNormals): a corner takes
 * the normalised sum of the facet normals of the vertex's triangles within
 * `crease_deg` of the facet normal of the vertex's last triangle in file
 * order, or its own facet normal when its triangle is not within it (the
 * reference uses 30 degrees, noor/source/objLoader.cpp:36-38).  A group whose
 * name contains "light" becomes a quad light from its first triangle
 * (noor/source/objLoader.cpp:56-79). */
noor_status noor_obj_load(const char* path, int unitize, float crease_deg, noor_obj** out);
noor_status noor_obj_view_get(const noor_obj* obj, noor_obj_view* view);
/* Message of the last failed noor_obj_load in this thread. */
const char* noor_loader_last_error(void);
void noor_obj_free(noor_obj* obj);

/* Reads an image as RGBA32F, row 0 first: Radiance RGBE (.hdr, flat or
 * run-length scanlines, -Y +X orientation), PFM (colour or grey, either byte
 * order, stored bottom row first), binary PNM (P5 grey, P6 colour, 8 or 16
 * bits: (v / max)^2.2), PNG (plain or Adam7 interlaced; every colour type
 * and bit depth; samples cut to 8 bits, then colour (v / 255)^2.2 and alpha v / 255,
 * as stb_image's loadf converts them), TGA (by the .tga extension;
 * uncompressed or run-length true colour, grey or colour-mapped; converted
 * the same way), OpenEXR (single-part scanline; R G B
 * A or Y channels in HALF or FLOAT; NONE, RLE, ZIPS or ZIP compression, as
 * tinyexr reads them), baseline JPEG (sequential Huffman, 8-bit, grey or
 * YCbCr with sampling factors 1 or 2, restart intervals; chroma upsampled by
 * replication; converted as PNG).  Grey is replicated to rgb; alpha is 1
 * where the file has none.  The returned buffer is freed with
 * noor_image_free. */
noor_status noor_image_load(const char* path, uint32_t* width, uint32_t* height, float** rgba);
void noor_image_free(float* rgba);

#ifdef __cplusplus
}
#endif
#endif
