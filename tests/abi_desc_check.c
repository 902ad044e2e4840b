/*
 * Fills lucida_scene_desc field by field the way the Rust binding's
 * scene_desc() fills it from a reference SceneDef (blob: analytic
 * primitives, instances, flattened BLAS + TLAS nodes, leaf-ordered
 * triangles, lights; tlas_root; the environment map), creates a scene from
 * it, renders it, and compares the film bit for bit with the same scene
 * uploaded through the library's own path. Two scenes: instance-forest
 * (instanced, lit by a procedural sky) and sunset (analytic, lit only by its
 * sky). Exit status 0 when every film matches.
 *
 * Plain C against include/lucida.h and include/lucida_scenes.h; built with
 * g++ by __graft_entry__.build(), run by tests/test_abi_desc_gpu.py.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lucida.h"
#include "lucida_scenes.h"

#define CHECK(x)                                                                              \
    do {                                                                                      \
        lucida_status s_ = (x);                                                               \
        if (s_ != LUCIDA_OK) {                                                                \
            fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #x, (int)s_,      \
                    lucida_last_error());                                                     \
            exit(2);                                                                          \
        }                                                                                     \
    } while (0)

static int render_and_compare(lucida_ctx* ctx, const char* label, lucida_scene* a, lucida_scene* b,
                              const GpuUniforms* u, uint32_t spp) {
    lucida_film *fa, *fb;
    CHECK(lucida_film_create(ctx, u->width, u->height, &fa));
    CHECK(lucida_film_create(ctx, u->width, u->height, &fb));
    CHECK(lucida_render_wavefront(ctx, a, u, spp, fa, NULL, NULL));
    CHECK(lucida_render_wavefront(ctx, b, u, spp, fb, NULL, NULL));
    size_t n = (size_t)u->width * u->height;
    GpuAccum* x = (GpuAccum*)malloc(n * sizeof(GpuAccum));
    GpuAccum* y = (GpuAccum*)malloc(n * sizeof(GpuAccum));
    CHECK(lucida_film_download(fa, x));
    CHECK(lucida_film_download(fb, y));
    double energy = 0.0;
    for (size_t i = 0; i < n; ++i) energy += x[i].radiance[0] / (x[i].samples > 0 ? x[i].samples : 1);
    int same = memcmp(x, y, n * sizeof(GpuAccum)) == 0;
    printf("%s: %s, mean red %.6f\n", label, same ? "films bit-identical" : "FILMS DIFFER", energy / (double)n);
    free(x);
    free(y);
    lucida_film_destroy(fa);
    lucida_film_destroy(fb);
    return same && energy > 0.0 ? 0 : 1;
}

int main(void) {
    lucida_ctx* ctx;
    CHECK(lucida_ctx_create(0, &ctx));
    int failures = 0;

    /* --- instance-forest: the flattened two-level blob ------------------- */
    {
        lucida_blob* blob;
        CHECK(lucida_blob_create("instance-forest", NULL, &blob));
        lucida_scene_desc own;
        CHECK(lucida_blob_desc(blob, &own));
        const lucida_mesh_range* meshes;
        const lucida_instance_placement* placements;
        size_t nm, np;
        CHECK(lucida_blob_instancing(blob, &meshes, &nm, &placements, &np));
        lucida_scene* built;
        CHECK(lucida_scene_create_instanced(ctx, &own, meshes, nm, placements, np, &built));
        /* what Scene::build_instances leaves in the reference's blob */
        size_t ntris, nnodes, ninst;
        uint32_t tlas_root;
        CHECK(lucida_scene_counts(built, &ntris, &nnodes, &ninst, &tlas_root));
        GpuBvhNode* nodes = (GpuBvhNode*)malloc(nnodes * sizeof(GpuBvhNode));
        GpuTriangle* tris = (GpuTriangle*)malloc(ntris * sizeof(GpuTriangle));
        GpuInstance* inst = (GpuInstance*)malloc(ninst * sizeof(GpuInstance));
        CHECK(lucida_scene_download_bvh(built, nodes, tris));
        CHECK(lucida_scene_download_instances(built, inst));

        /* scene_desc(): every field, as the binding fills it */
        lucida_scene_desc d;
        memset(&d, 0, sizeof d); /* the binding's manual Default: null pointers, zero counts */
        d.materials = own.materials;
        d.num_materials = own.num_materials;
        d.primitives = own.primitives;       /* analytic primitives only */
        d.num_primitives = own.num_primitives;
        d.instances = inst;                  /* blob.instances */
        d.num_instances = ninst;
        d.positions = own.positions;
        d.num_positions = own.num_positions;
        d.vertex_attrs = own.vertex_attrs;
        d.triangles = tris;                  /* BLAS leaf order */
        d.num_triangles = ntris;
        d.bvh_nodes = nodes;                 /* BLAS nodes, then the TLAS */
        d.num_bvh_nodes = nnodes;
        d.lights = own.lights;
        d.num_lights = own.num_lights;
        d.tlas_root = tlas_root;
        d.env_width = own.env_width;
        d.env_height = own.env_height;
        d.env_pixels = own.env_pixels;
        d.env_conditional = own.env_conditional;
        d.env_marginal = own.env_marginal;
        d.env_total_weight = own.env_total_weight;
        d.build_lbvh = 0;
        lucida_scene* via_desc;
        CHECK(lucida_scene_create(ctx, &d, &via_desc));
        GpuUniforms u;
        CHECK(lucida_blob_uniforms(blob, 96, 64, &u));
        u.samples_per_launch = 8;
        u.max_depth = 6;
        if (u.num_instances != 0 || ninst != 49 || d.env_width != 256) {
            fprintf(stderr, "instance-forest: unexpected counts (%zu instances, env %u)\n", ninst, d.env_width);
            failures++;
        }
        u.num_instances = (uint32_t)ninst;
        u.tlas_root = tlas_root;
        u.num_bvh_nodes = (uint32_t)nnodes;
        failures += render_and_compare(ctx, "instance-forest via scene_desc()", via_desc, built, &u, 8);
        lucida_scene_destroy(via_desc);
        lucida_scene_destroy(built);
        free(nodes);
        free(tris);
        free(inst);
        lucida_blob_destroy(blob);
    }

    /* --- sunset: analytic primitives under the sky ------------------------ */
    {
        lucida_blob* blob;
        CHECK(lucida_blob_create("sunset", NULL, &blob));
        lucida_scene_desc own;
        CHECK(lucida_blob_desc(blob, &own));
        lucida_scene* built;
        CHECK(lucida_scene_create(ctx, &own, &built));
        lucida_scene_desc d;
        memset(&d, 0, sizeof d);
        d.materials = own.materials;
        d.num_materials = own.num_materials;
        d.primitives = own.primitives;
        d.num_primitives = own.num_primitives;
        d.instances = own.instances; /* an empty Vec's pointer, count 0 */
        d.num_instances = 0;
        d.positions = own.positions;
        d.num_positions = own.num_positions;
        d.vertex_attrs = own.vertex_attrs;
        d.triangles = own.triangles;
        d.num_triangles = own.num_triangles;
        d.bvh_nodes = own.bvh_nodes;
        d.num_bvh_nodes = own.num_bvh_nodes;
        d.lights = own.lights;
        d.num_lights = own.num_lights;
        d.tlas_root = 0;
        d.env_width = own.env_width;
        d.env_height = own.env_height;
        d.env_pixels = own.env_pixels;
        d.env_conditional = own.env_conditional;
        d.env_marginal = own.env_marginal;
        d.env_total_weight = own.env_total_weight;
        lucida_scene* via_desc;
        CHECK(lucida_scene_create(ctx, &d, &via_desc));
        GpuUniforms u;
        CHECK(lucida_blob_uniforms(blob, 96, 64, &u));
        u.samples_per_launch = 8;
        u.max_depth = 6;
        if (d.env_width != 256 || d.env_height != 128 || d.num_lights != 0) {
            fprintf(stderr, "sunset: unexpected environment\n");
            failures++;
        }
        failures += render_and_compare(ctx, "sunset via scene_desc()", via_desc, built, &u, 8);
        lucida_scene_destroy(via_desc);
        lucida_scene_destroy(built);
        lucida_blob_destroy(blob);
    }

    lucida_ctx_destroy(ctx);
    printf("%s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
