/* Field offsets of the boundary PODs against the reference's CUDA structs.
 *
 * Each expected offset is derived from the reference declaration cited
 * (CUDA vector types: float2 is 8 bytes, float3 12 bytes aligned to 4; uint
 * and enums 4 bytes; unsigned short 2; bool 1), so that a reference host can
 * hand its structs over byte for byte.  Compiled by tests/test_abi.py with
 * gcc -fsyntax-only: a wrong offset fails the compile. */
#include <stddef.h>

#include "noor_hip.h"

#define AT(type, field, off) _Static_assert(offsetof(type, field) == (off), #type "." #field " at " #off)

/* CudaMaterial (noor/cuda/material.cuh:26-58): 7 float2, 10 uint, 7 float, uint */
AT(noor_material, diffuse_uvscale, 0);
AT(noor_material, specular_uvscale, 8);
AT(noor_material, transparency_uvscale, 16);
AT(noor_material, emittance_uvscale, 24);
AT(noor_material, bump_uvscale, 32);
AT(noor_material, roughness_uvscale, 40);
AT(noor_material, metalness_uvscale, 48);
AT(noor_material, diffuse_tex_idx, 56);
AT(noor_material, specular_tex_idx, 60);
AT(noor_material, transmission_tex_idx, 64);
AT(noor_material, emittance_tex_idx, 68);
AT(noor_material, bump_tex_idx, 72);
AT(noor_material, transparency_tex_idx, 76);
AT(noor_material, ior_tex_idx, 80);
AT(noor_material, k_tex_idx, 84);
AT(noor_material, roughness_tex_idx, 88);
AT(noor_material, metalness_tex_idx, 92);
AT(noor_material, bumpfactor, 96);
AT(noor_material, emitter_scale, 100);
AT(noor_material, coat_weight, 104);
AT(noor_material, coat_roughness, 108);
AT(noor_material, coat_thickness, 112);
AT(noor_material, coat_sigma, 116);
AT(noor_material, coat_ior, 120);
AT(noor_material, type, 124);
_Static_assert(sizeof(noor_material) == 128, "noor_material size");

/* CudaArHosekSkyModelState (noor/cuda/hosek.cuh:30-40): float[3][9],
   noor_float3, 2 float, 2 noor_float3, 2 float */
AT(noor_hosek_state, configs, 0);
AT(noor_hosek_state, radiances, 108);
AT(noor_hosek_state, turbidity, 120);
AT(noor_hosek_state, solar_radius, 124);
AT(noor_hosek_state, emission_correction_factor_sky, 128);
AT(noor_hosek_state, emission_correction_factor_sun, 140);
AT(noor_hosek_state, albedo, 152);
AT(noor_hosek_state, elevation, 156);
_Static_assert(sizeof(noor_hosek_state) == 160, "noor_hosek_state size");
/* CudaHosekSky (noor/cuda/hosek.cuh:43-51): 3 states, 2 float3, 2 float */
AT(noor_hosek, state_r, 0);
AT(noor_hosek, state_g, 160);
AT(noor_hosek, state_b, 320);
AT(noor_hosek, solar_dir, 480);
AT(noor_hosek, ground_albedo, 492);
AT(noor_hosek, solar_scale, 504);
AT(noor_hosek, solar_radius, 508);
_Static_assert(sizeof(noor_hosek) == 512, "noor_hosek size");

/* CudaSpec (noor/cuda/spec.cuh:41-58): 2 float3, 4 float, uint, 5 ushort,
   the SkydomeType enum (4-aligned), 3 bool */
AT(noor_spec, white, 0);
AT(noor_spec, black, 12);
AT(noor_spec, reflection_bias, 24);
AT(noor_spec, shadow_bias, 28);
AT(noor_spec, world_radius, 32);
AT(noor_spec, wr2, 36);
AT(noor_spec, bvh_root_node, 40);
AT(noor_spec, lighting_type, 44);
AT(noor_spec, bounces, 46);
AT(noor_spec, rr, 48);
AT(noor_spec, bvh_height, 50);
AT(noor_spec, num_gpus, 52);
AT(noor_spec, skydome_type, 56);
AT(noor_spec, debug_sky, 60);
AT(noor_spec, mis, 61);
AT(noor_spec, outofsync, 62);
_Static_assert(sizeof(noor_spec) == 64, "noor_spec size");

/* CudaTransform (noor/cuda/transform.cuh:26-28): Matrix3x3 of three
   noor_float3 rows (noor/cuda/math.cuh:150-153), then a float3 */
AT(noor_transform, row, 0);
AT(noor_transform, t, 36);
_Static_assert(sizeof(noor_transform) == 48, "noor_transform size");
/* CudaCamera (noor/cuda/camera.cuh:95-102): 3 transforms, 2 float3, 3 float,
   3 uint, the CameraType enum */
AT(noor_camera, camera_to_world, 0);
AT(noor_camera, raster_to_camera, 48);
AT(noor_camera, world_to_raster, 96);
AT(noor_camera, dx_camera, 144);
AT(noor_camera, dy_camera, 156);
AT(noor_camera, lens_radius, 168);
AT(noor_camera, focal_length, 172);
AT(noor_camera, image_area, 176);
AT(noor_camera, w, 180);
AT(noor_camera, h, 184);
AT(noor_camera, center, 188);
AT(noor_camera, type, 192);
_Static_assert(sizeof(noor_camera) == 196, "noor_camera size");

int main(void) { return 0; }
