/*
 * oracle.h -- CPU restatement of nvpro-pipeline's scene-update path
 * (transform propagation + bounding-box frustum culling).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is the checker that the HIP kernels are compared
 * against in tests/, in __graft_entry__.smoke() and in bench.py's CPU
 * baseline leg.
 *
 * Build: gcc -O2 -ffp-contract=off (no -ffast-math, no FMA).  Every FP32
 * operation is a separate IEEE add or multiply evaluated in the order the
 * reference's C++ source evaluates it (the reference builds without FMA and
 * without fast-math).
 *
 * Conventions (identical to the reference's dp::math types):
 *   - a matrix is 16 floats, row-major (m[r*4+c] = row r, column c);
 *   - vectors are row vectors, v' = v * M;
 *   - an object's box is stored the way ManagerBitSet::objectSetBoundingBox
 *     stores it: lowerLeft = (lower.xyz, 1), extent = (upper-lower, 0)
 *     (dp/culling/src/ManagerBitSet.cpp:99-109);
 *   - visibility masks are dp::util::BitArray words: bit i of the mask is
 *     bit (i & 63) of 64-bit word i >> 6, read back as little-endian uint32
 *     words (dp/util/BitArray.h:215-233, dp/culling/cpu/src/ManagerImpl.cpp:517).
 *
 * How each function is pinned (checked against something other than a
 * retyped copy of its own formula) -- see tests/test_oracle_*.py:
 *   orc_vec4_mul_mat44, orc_mat44_mul       bit-exact vs the compiled reference
 *                                           dp/math/Matmnt.h operators (oracle/_ref)
 *                                           and the committed golden vectors.
 *   orc_tree_compute                        bit-exact vs the compiled reference
 *                                           dp::transform::Tree::compute (oracle/_ref)
 *                                           on golden traces (add/remove/update).
 *   orc_obb_build, orc_is_visible_obb,
 *   orc_cull_group, orc_bbox_group          the reference's culling TUs need Boost
 *                                           headers this image lacks, so they are
 *                                           not compiled.  Pinned instead by the
 *                                           reference operators above (each step is
 *                                           one of them), hand-derived known-answer
 *                                           cases that separate every association
 *                                           choice (two-stage OBB rounding, corner
 *                                           add tree, <= vs <, else-if flags, NaN),
 *                                           and an FP64 brute-force check away from
 *                                           plane boundaries.
 *   orc_result_diff                         brute force (per-bit loop) and XOR
 *                                           invariants.
 *   orc_invert_transpose                    bit-exact vs the compiled reference
 *                                           dp::math::invertTranspose (oracle/_ref)
 *                                           on committed golden vectors, plus
 *                                           M^T x result = I in FP64 and the
 *                                           singular (det == 0) cases.
 */
#ifndef DPCULL_ORACLE_H
#define DPCULL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dp/math/Matmnt.h:1371-1379: out[c] = ((v0*m0c + v1*m1c) + v2*m2c) + v3*m3c. */
void orc_vec4_mul_mat44(const float v[4], const float m[16], float out[4]);

/* dp/math/Matmnt.h:1381-1417: out[i][j] = ((a_i0*b_0j + a_i1*b_1j) + a_i2*b_2j) + a_i3*b_3j. */
void orc_mat44_mul(const float a[16], const float b[16], float out[16]);

/* Same association in double precision (the FP64 oracle for error reports). */
void orc_mat44_mul_f64(const double a[16], const double b[16], double out[16]);

/* GroupCPU::updateOBBs scalar branch (dp/culling/cpu/src/ManagerImpl.cpp:125-156):
 *   obb[0..3]  = point = lowerLeft * M
 *   obb[4..7]  = ex    = extent[0] * M[0]
 *   obb[8..11] = ey    = extent[1] * M[1]
 *   obb[12..15]= ez    = extent[2] * M[2]                                   */
void orc_obb_build(const float lower4[4], const float extent4[4], const float m[16], float obb[16]);

/* GroupCPU::updateOBBs over a whole group (ManagerImpl.cpp:114-162): obbs = n x 16 floats. */
void orc_obb_build_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                         const void* matrices, size_t stride, float* obbs);

/* isVisible(projection, OBB) + determineCullFlags
 * (dp/culling/cpu/src/ManagerImpl.cpp:199-229, 263-289).  Returns 1 if visible. */
int orc_is_visible_obb(const float vp[16], const float obb[16]);

/* Per-corner clip coordinates in the reference's add tree (ManagerImpl.cpp:269-281);
 * corners[k*4 .. k*4+3] = vectors[k]. Exposed for known-answer tests. */
void orc_obb_corners_clip(const float vp[16], const float obb[16], float corners[32]);

/* cpu::ManagerImpl::cull on one group (ManagerImpl.cpp:467-518, scalar path):
 * builds every OBB (updateOBBs with a dirty cache), tests each, writes the
 * BitArray words.  mask must hold 2*ceil(n/64) uint32 words; bits >= n are 0.
 *   lower4/extent4 : n x 4 floats (AoS, as ObjectBitSet stores them)
 *   xform          : n transform indices (ObjectBitSet::m_transformIndex)
 *   matrices       : base pointer, stride in bytes (GroupBitSet::setMatrices) */
void orc_cull_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                    const void* matrices, size_t stride, const float vp[16], uint32_t* mask);

/* Same loop over precomputed OBBs (n x 16 floats). */
void orc_cull_obbs(size_t n, const float* obbs, const float vp[16], uint32_t* mask);

/* ManagerBitSet::calculateBoundingBox scalar branch (ManagerBitSet.cpp:268-306) and the
 * Box3f(min, max) constructor it returns through (dp/math/Boxnt.h:213-258).
 * out6 = lower.xyz, upper.xyz.  An empty group yields (-FLT_MAX, +FLT_MAX). */
void orc_bbox_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                    const void* matrices, size_t stride, float out6[6]);

/* dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166).
 * entries: pairs (parent, transform) level by level, level l occupying
 * [level_offsets[l], level_offsets[l+1]).  local/world: AoS 16 floats per index.
 * dirty_local / dirty_world: one byte per index (m_dirtyTransforms / m_dirtyWorldMatrices).
 * On return dirty_world_event (if not NULL) holds the bits the
 * EventWorldMatricesChanged carried, and both dirty arrays are cleared. */
void orc_tree_compute(size_t n_levels, const uint64_t* level_offsets, const uint32_t* entries,
                      const float* local, float* world, size_t n_index,
                      uint8_t* dirty_local, uint8_t* dirty_world, uint8_t* dirty_world_event);

/* FP64 evaluation of the same recurrence (all entries recomputed). */
void orc_tree_compute_f64(size_t n_levels, const uint64_t* level_offsets, const uint32_t* entries,
                          const double* local, double* world);

/* ResultBitSet::updateChanged diff step (dp/culling/src/ResultBitSet.cpp:100-107):
 * changed = new ^ old over n bits; writes the changed indices in ascending
 * order (BitArray::traverseBits order) and returns their count. */
size_t orc_result_diff(size_t n, const uint32_t* old_mask, const uint32_t* new_mask, uint32_t* changed_idx);

/* dp::math::invertTranspose (dp/math/Matmnt.h:1171-1201): the cofactor form of
 * the inverse, written transposed.  Returns 1 and writes out when det != 0
 * (NaN included); returns 0 and leaves out untouched when det == 0. */
int orc_invert_transpose(const float m[16], float out[16]);

/* orc_obb_build_group and orc_cull_obbs over `threads` host threads (a static
 * partition of the objects into whole 64-bit mask words); same results.  For
 * timing the CPU path on every core. */
void orc_obb_build_group_mt(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                            const void* matrices, size_t stride, float* obbs, int threads);
void orc_cull_obbs_mt(size_t n, const float* obbs, const float vp[16], uint32_t* mask, int threads);

#ifdef __cplusplus
}
#endif

#endif
