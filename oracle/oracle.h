/*
 * oracle.h — C entry points of the CPU oracle, for the test suite (ctypes).
 *
 * TEST INFRASTRUCTURE. The oracle is a restatement, in C++ with contraction
 * off, of the reference's Rust CPU twin (crates/core). Only tests, the smoke
 * check and the bench's CPU-baseline leg load it; the product never does.
 */
#ifndef LUCIDA_ORACLE_H
#define LUCIDA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/lucida.h"

#ifdef __cplusplus
extern "C" {
#endif

/* lbvh.rs:72 / :92 / :109 / :149 */
uint32_t oracle_expand_bits(uint32_t v);
uint32_t oracle_morton3d(float x, float y, float z);
void oracle_normalise(const float p[3], const float lo[3], const float hi[3], float out[3]);
int oracle_lcp(const uint32_t* codes, size_t n, int i, int j);
/* bvh.rs:202 */
void oracle_triangle_bounds(const GpuTriangle* tris, size_t n, const float* positions, LucidaAabb* out);

/* Lbvh::build_with_leaf_size (lbvh.rs:422) with every stage exposed. Arrays
 * are caller-owned: codes/order [n], karras [2(n-1)], parent [2n-1],
 * node_bounds [2n-1], subtree [n-1], nodes [2n-1], prim_indices [n]. Any may be
 * NULL. Returns the node count. */
size_t oracle_lbvh_build(const GpuTriangle* tris, size_t n, const float* positions, size_t max_leaf,
                         uint32_t* codes, uint32_t* order, uint32_t* karras, uint32_t* parent,
                         LucidaAabb* node_bounds, uint32_t* subtree, LucidaAabb* scene_bounds, GpuBvhNode* nodes,
                         uint32_t* prim_indices, float* sah_cost);

/* bvh.rs:469 / :544 (returns the number of problems found). */
float oracle_sah_cost(const GpuBvhNode* nodes, size_t n);
size_t oracle_validate(const GpuBvhNode* nodes, size_t n, const uint32_t* prim_indices, size_t num_prims,
                       size_t triangle_count);

/* Treelet restructuring of the LBVH (treelet.cpp; no reference counterpart):
 * the CPU statement of lucida_bvh_build_treelet. nodes [2n-1], prim_indices
 * [n]; returns the node count. _with takes the cost model's constants. */
size_t oracle_treelet_build(const GpuTriangle* tris, size_t n, const float* positions, uint32_t rounds,
                            GpuBvhNode* nodes, uint32_t* prim_indices, float* sah_cost);
size_t oracle_treelet_build_with(const GpuTriangle* tris, size_t n, const float* positions, uint32_t rounds,
                                 float ci, float ct, uint32_t max_leaf, uint32_t min_size, GpuBvhNode* nodes,
                                 uint32_t* prim_indices, float* sah_cost);
/* The subset search over k <= 7 leaves: least cost, best split's left side
 * and leaf choice per subset mask ([2^k] each; entry 0 unused). */
void oracle_treelet_dp(int k, const LucidaAabb* leaf_box, const float* leaf_cost, const uint32_t* leaf_size,
                       float* cost_out, uint8_t* part_out, uint8_t* leafy_out);

/* Stable (key, value) sort on the CPU: the slice::sort of (key, index) pairs
 * the reference's radix tests compare against (crates/gpu/tests/radix_sort.rs). */
void oracle_sort_pairs(const uint32_t* keys, const uint32_t* vals, size_t n, uint32_t* keys_out, uint32_t* vals_out);

#ifdef __cplusplus
}
#endif
#endif
